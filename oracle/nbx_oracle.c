/*
 * CPU oracle for the NBNxM short-range nonbonded path (LJ + Ewald real space).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference's algorithm, used as the checker for the HIP product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library never links or calls it.
 *
 * It is compiled twice: with ORC_DOUBLE undefined ("real" = float, the
 * reference's mixed-precision build) and with ORC_DOUBLE defined ("real" =
 * double, the reference's double-precision build that the FP64 oracle stands
 * for).  Build with -ffp-contract=off so every float operation is rounded as
 * written; the canonical squared distance used for the pair set is
 * ((dx*dx + dy*dy) + dz*dz) with no fused multiply-add.
 *
 * Reference files restated here (paths relative to the reference root):
 *   src/gromacs/ewald/ewald_utils.cpp:46-74           calc_ewaldcoeff_q
 *   src/gromacs/tables/forcetable.cpp:114-387         Ewald correction tables
 *   src/gromacs/mdlib/forcerec.cpp:680-687,726-764    sh_ewald, table setup
 *   api/nblib/gmxsetup.cpp:143-166,226-284            nbfp, interaction constants
 *   src/gromacs/nbnxm/grid.cpp:103-262,292-427,454-484,860-1048,1173-1445
 *   src/gromacs/nbnxm/atomdata.cpp:134-250,440-620,723-836,1018-1073,1181-1228
 *   src/gromacs/nbnxm/pairlist.cpp:255-372,907-1078,1367-1435,2002-2067,
 *                                  2241-2322,2841-2878,3076-3582
 *   src/gromacs/nbnxm/kernels_reference/kernel_ref_outer.h:85-409,
 *                                       kernel_ref_inner.h:44-375
 *   src/gromacs/pbcutil/pbc.cpp:1187-1202, pbcutil/ishift.h:40-55
 *   src/gromacs/mdlib/calcvir.cpp:103-148             virial
 *   src/gromacs/mdlib/forcerec.cpp:703-717,780-818,848-873  LJ modifiers, LJ-PME
 *   src/gromacs/ewald/ewald_utils.cpp:76-115          calc_ewaldcoeff_lj
 *   src/gromacs/nbnxm/atomdata.cpp:288-330            LJ-PME grid parameters
 *   src/gromacs/nbnxm/kernels_reference/kernel_ref_prune.cpp:39-123  pruning
 *
 * Parity pins (see tests/test_oracle_*.py): nblib's Argon and SPC-methanol
 * golden forces (RF/cut-off path), closed forms of the Ewald correction,
 * brute-force O(N^2) pair sets and FP64 pair sums, central finite differences
 * of the energy, Newton's third law and the virial closed form; for the LJ
 * switches and LJ-PME, energy gradients against forces, zero potential and
 * force at the cut-off, the plain LJ below the switch radius, the
 * smoothstep's value at mid-range, and LJ-PME's two-atom closed form, tail
 * condition and self-energy limit.
 */
#include <float.h>
#include <math.h>
#ifdef ORC_OMP
#    include <omp.h>
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORC_DOUBLE
typedef double real;
#    define R_SQRT sqrt
#    define R_ERFC erfc
#    define R_CBRT cbrt
#    define R_REAL_EPS 2.2204460492503131e-16
#    define R_REAL_MAX DBL_MAX
#    define ORC_MIN_RSQ 1.0e-36
#else
typedef float real;
#    define R_SQRT sqrtf
#    define R_ERFC erfcf
#    define R_CBRT cbrtf
#    define R_REAL_EPS 1.19209290e-07F
#    define R_REAL_MAX FLT_MAX
#    define ORC_MIN_RSQ 3.82e-07F
#endif

#define ORC_FLOAT_EPS 1.19209290e-07F
#define ORC_NC 4 /* atoms per i- and j-cluster (plain-C 4x4 list) */
#define ORC_SHIFTS 45
#define ORC_CENTRAL 22
#define ORC_MASK_ALL 0xffffffffU
#define ORC_MASK_DIAG 0x08ceU

/* XYZ2IS from pbcutil/ishift.h:47-50: D_BOX_X = 2, D_BOX_Y = D_BOX_Z = 1 */
static int xyz2is(int x, int y, int z)
{
    return 5 * (3 * (z + 1) + y + 1) + x + 2;
}

/* ONE_4PI_EPS0 from math/units.h:57-79 */
double orc_one_4pi_eps0(void)
{
    const double E_CHARGE    = 1.602176634e-19;
    const double AVOGADRO    = 6.02214076e23;
    const double EPSILON0_SI = 8.8541878128e-12;
    const double EPSILON0    = (EPSILON0_SI * 1e-9 * 1e3) / (E_CHARGE * E_CHARGE * AVOGADRO);
    return 1.0 / (4.0 * M_PI * EPSILON0);
}

int orc_real_size(void)
{
    return (int)sizeof(real);
}

/* ------------------------------------------------------------------ */
/* Ewald coefficient and correction tables                             */
/* ------------------------------------------------------------------ */

/* calc_ewaldcoeff_q, ewald/ewald_utils.cpp:46-74: doubling then bisection
 * on erfc(beta*rc) = rtol, all in "real". */
real orc_ewald_coeff(real rc, real rtol)
{
    real beta = 5, low, high;
    int  n, i = 0;
    do
    {
        i++;
        beta *= 2;
    } while (R_ERFC(beta * rc) > rtol);
    n    = i + 60;
    low  = 0;
    high = beta;
    for (i = 0; i < n; i++)
    {
        beta = (low + high) / 2;
        if (R_ERFC(beta * rc) > rtol)
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

/* spline3_table_scale, tables/forcetable.cpp:304-316 */
static double spline3_table_scale(double third_deriv_max, double x_scale, double func_tol)
{
    double deriv_tol = ORC_FLOAT_EPS;
    double sc_deriv  = sqrt(third_deriv_max / (6 * 4 * deriv_tol * x_scale)) * x_scale;
    if (func_tol < (double)R_REAL_EPS)
    {
        func_tol = (double)R_REAL_EPS;
    }
    double sc_func = cbrt(third_deriv_max / (6 * 12 * sqrt(3.0) * func_tol)) * x_scale;
    return sc_deriv > sc_func ? sc_deriv : sc_func;
}

/* ewald_spline3_table_scale (Coulomb part), tables/forcetable.cpp:331-358 */
real orc_ewald_table_scale(real beta, real rc)
{
    double erf_x_d3 = 1.0522;
    double etol     = 0.1 * R_ERFC(beta * rc);
    real   sc_q     = (real)spline3_table_scale(erf_x_d3, beta, etol);
    real   sc       = 0;
    return sc_q > sc ? sc_q : sc;
}

/* Table size, mdlib/forcerec.cpp:755 */
int orc_ewald_table_size(real rc, real scale)
{
    return (int)(rc * scale) + 2;
}

/* v_q_ewald_lr, tables/forcetable.cpp:114-124 */
static double v_q_ewald_lr(double beta, double r)
{
    if (r == 0)
    {
        return beta * 2 / sqrt(M_PI);
    }
    return erf(beta * r) / r;
}

/* generateEwaldCorrectionTables, tables/forcetable.cpp:144-298.
 * F, V have n entries, FDV0 has 4*n. */
void orc_ewald_tables(int numPoints, double tableScaling, real beta, real* table_f, real* table_v, real* table_fdv0)
{
    real   tab_max;
    int    i, i_inrange;
    double dc, dc_new;
    int    bOutOfRange;
    double v_r0, v_r1, v_inrange, vi, a0, a1, a2dx;
    double x_r0;

    const double dx = 1 / tableScaling;

    for (i = 0; i < numPoints; i++)
    {
        table_f[i] = 0;
        table_v[i] = 0;
    }
    tab_max     = (real)(R_REAL_MAX * 0.0001);
    bOutOfRange = 0;
    i_inrange   = numPoints;
    v_inrange   = 0;
    dc          = 0;
    for (i = numPoints - 1; i >= 0; i--)
    {
        x_r0 = i * dx;
        v_r0 = v_q_ewald_lr(beta, x_r0);
        if (!bOutOfRange)
        {
            i_inrange = i;
            v_inrange = v_r0;
            vi        = v_r0;
        }
        else
        {
            vi = v_inrange - dc * (i - i_inrange) * dx;
        }
        table_v[i] = (real)vi;
        if (i == 0)
        {
            continue;
        }
        v_r1 = v_q_ewald_lr(beta, (i - 1) * dx);
        if (v_r1 != v_r1 || v_r1 < -tab_max || v_r1 > tab_max)
        {
            bOutOfRange = 1;
        }
        if (!bOutOfRange)
        {
            a2dx = (v_r0 + v_r1 - 2 * v_q_ewald_lr(beta, x_r0 - 0.5 * dx)) / (0.25 * dx);
            dc   = (v_r0 - v_r1) / dx + 0.5 * a2dx;
        }
        if (i == numPoints - 1)
        {
            table_f[i] = (real)(-dc);
        }
        else
        {
            table_f[i] += -0.5 * dc;
        }
        if (!bOutOfRange)
        {
            a0     = v_r0;
            a1     = dc;
            a2dx   = (a1 * dx + v_r1 - a0) * 2 / dx;
            dc_new = a1 - a2dx;
            if (dc_new != dc_new || dc_new < -tab_max || dc_new > tab_max)
            {
                bOutOfRange = 1;
            }
            else
            {
                dc = dc_new;
            }
        }
        table_f[(i - 1)] = (real)(-0.5 * dc);
    }
    table_f[0] *= 2;

    for (i = 0; i < numPoints - 1; i++)
    {
        table_fdv0[4 * i]     = table_f[i];
        table_fdv0[4 * i + 1] = table_f[i + 1] - table_f[i];
        table_fdv0[4 * i + 2] = table_v[i];
        table_fdv0[4 * i + 3] = 0;
    }
    const int last            = numPoints - 1;
    table_fdv0[4 * last]      = table_f[last];
    table_fdv0[4 * last + 1]  = -table_f[last];
    table_fdv0[4 * last + 2]  = table_v[last];
    table_fdv0[4 * last + 3]  = 0;
}

/* Exact closed forms of the Ewald correction, in double.
 * V(r) = erf(beta r)/r;  F(r) = -dV/dr = erf(beta r)/r^2 - 2 beta/sqrt(pi) exp(-beta^2 r^2)/r.
 * At r = 0 the limits are V = 2 beta/sqrt(pi), F = 0. */
double orc_ewald_v_exact(double beta, double r)
{
    return v_q_ewald_lr(beta, r);
}

double orc_ewald_f_exact(double beta, double r)
{
    if (r == 0)
    {
        return 0;
    }
    double br = beta * r;
    if (br < 1e-3)
    {
        /* Taylor series of -d/dr erf(beta r)/r:
         * F = 4 beta^3 r / (3 sqrt(pi)) * (1 - 3/5 z^2 + 3/14 z^4), z = beta r */
        double z2 = br * br;
        return 4.0 * beta * beta * beta * r / (3.0 * sqrt(M_PI)) * (1 - 0.6 * z2 + 3.0 / 14.0 * z2 * z2);
    }
    return erf(br) / (r * r) - 2 * beta / sqrt(M_PI) * exp(-br * br) / r;
}

/* ------------------------------------------------------------------ */
/* System state                                                        */
/* ------------------------------------------------------------------ */

typedef struct
{
    /* interaction constants (api/nblib/gmxsetup.cpp:226-284) */
    int    coulomb; /* 0 = reaction-field / plain cut-off, 1 = Ewald */
    int    ewald_exact; /* 0 = reference spline table, 1 = exact closed form */
    real   rlist, rcoulomb;
    /* interaction_const_t::rvdw: rcoulomb unless a twin cut-off is set */
    real   rvdw;
    real   epsfac, k_rf, c_rf, beta, sh_ewald, cpot_disp, cpot_rep;
    /* LJ modifier (interaction_const_t vdw_modifier): 0 potential shift,
     * 1 force switch, 2 potential switch; rvdw_switch and the constants of
     * forcerec.cpp:780-818 (dispersion_shift, repulsion_shift, vdw_switch) */
    int    vdw_mod;
    real   rvdw_switch;
    real   fsw_c2d, fsw_c3d, fsw_c2r, fsw_c3r;
    real   sw_c3, sw_c4, sw_c5;
    /* LJ-PME real space (vdwtype evdwPME): grid combination rule 0
     * geometric, 1 Lorentz-Berthelot; beta_lj^2, beta_lj^6 / 6, sh_lj_ewald */
    int    lj_pme, ljpme_comb;
    real   lje_coeff2, lje_coeff6_6, lje_vc;
    real   tab_scale;
    int    tab_n;
    real*  tabF;
    real*  tabV;
    real*  tabFDV0;
    /* LJ parameters (6*C6, 12*C12), numTypes includes the zero padding type */
    int   numTypes;
    real* nbfp;
    /* exclusions, CSR with self included */
    int  natoms;
    int* excl_off;
    int* excl_idx;
    /* box and shifts */
    real box[3];
    real shiftvec[ORC_SHIFTS][3];
    /* grid (grid.cpp) */
    real  density;
    int   ncx, ncy, ncol;
    real  cellSize[2], invCellSize[2];
    int*  cxy_na;
    int*  cxy_ind;
    int   ncells;
    int*  atomIndices; /* ncells*4, -1 = filler */
    int*  cells;       /* natoms: atom -> slot */
    real* xnb;         /* ncells*4*3 */
    float* bb;         /* ncells*6 lower xyz, upper xyz */
    int*   nbtype;
    real*  nbq;
    /* pairlist (pairlist.h NbnxnPairlistCpu) */
    int       nci, ncj, cap_ci, cap_cj;
    int*      ci; /* 4 per entry: ci, shift, cj_ind_start, cj_ind_end */
    int*      cj;
    unsigned* excl;
    /* the list as searched (at rlist), kept by the first prune after a
     * search so that every prune starts from it (nbl->ciOuter/cjOuter) */
    int       haveOuter, nciO, ncjO;
    int*      ciO;
    int*      cjO;
    unsigned* exclO;
    /* outputs */
    real*  fnb; /* ncells*4*3 */
    real   fshift[ORC_SHIFTS * 3];
    real   Vvdw, Vc;
} orc_sys;

static void* xrealloc(void* p, size_t n)
{
    void* q = realloc(p, n ? n : 1);
    if (!q)
    {
        abort();
    }
    return q;
}

/* Interaction constants as nblib sets them (gmxsetup.cpp:226-284), with two
 * extensions: rcoulomb may be shorter than rlist, and sh_ewald may follow
 * mdrun's potential-shift (forcerec.cpp:680-687) instead of nblib's 0. */
orc_sys* orc_create(int coulomb, int ewald_exact, double rlist, double rcoulomb, double epsilon_r,
                    double epsilon_rf, double ewald_rtol, int ewald_potential_shift)
{
    orc_sys* s     = (orc_sys*)calloc(1, sizeof(orc_sys));
    s->coulomb     = coulomb;
    s->ewald_exact = ewald_exact;
    s->rlist       = (real)rlist;
    s->rcoulomb    = (real)rcoulomb;
    s->rvdw        = s->rcoulomb;
    real rvdw      = s->rvdw;
    /* power6/power12 of math/functions.h:333-390 */
    real p3        = rvdw * (rvdw * rvdw);
    real p6        = p3 * p3;
    real p12       = p6 * p6;
    s->cpot_disp   = (real)(-1.0 / p6);
    s->cpot_rep    = (real)(-1.0 / p12);
    s->epsfac      = (real)(orc_one_4pi_eps0() / (real)epsilon_r);
    {
        /* calc_rffac, mdlib/rf_util.cpp:54-66 */
        real eps_r = (real)epsilon_r, eps_rf = (real)epsilon_rf, Rc = s->rcoulomb;
        if (eps_rf == 0)
        {
            s->k_rf = 1 / (2 * Rc * Rc * Rc);
        }
        else
        {
            s->k_rf = (eps_rf - eps_r) / (2 * eps_rf + eps_r) / (Rc * Rc * Rc);
        }
        s->c_rf = 1 / Rc + s->k_rf * Rc * Rc;
    }
    if (coulomb == 1)
    {
        s->beta      = orc_ewald_coeff(s->rcoulomb, (real)ewald_rtol);
        s->sh_ewald  = ewald_potential_shift ? (real)(R_ERFC(s->beta * s->rcoulomb) / s->rcoulomb) : 0;
        s->tab_scale = orc_ewald_table_scale(s->beta, s->rcoulomb);
        s->tab_n     = orc_ewald_table_size(s->rcoulomb, s->tab_scale);
        s->tabF      = (real*)xrealloc(NULL, sizeof(real) * s->tab_n);
        s->tabV      = (real*)xrealloc(NULL, sizeof(real) * s->tab_n);
        s->tabFDV0   = (real*)xrealloc(NULL, sizeof(real) * 4 * s->tab_n);
        orc_ewald_tables(s->tab_n, s->tab_scale, s->beta, s->tabF, s->tabV, s->tabFDV0);
    }
    return s;
}

/* The twin cut-off: interaction_const_t::rvdw below rcoulomb
 * (kerneldispatch.cpp:175-192 picks the TAB_TWIN / EWALD_TWIN kernels, which
 * mask LJ beyond rvdw: VDW_CUTOFF_CHECK, kernel_ref_inner.h:254-271), with
 * the LJ potential shift at rvdw (forcerec.cpp:848-873).  grompp allows it
 * with Ewald electrostatics and an LJ cut-off only (readir.cpp:386-398).
 * Call before orc_set_vdw_modifier, which then switches at rvdw.  Returns
 * -1 unless 0 < rvdw <= rcoulomb with Ewald and no LJ-PME. */
int orc_set_vdw_cutoff(orc_sys* s, double rvdw)
{
    real r = (real)rvdw;
    if (!(r > 0 && r <= s->rcoulomb) || (r < s->rcoulomb && s->coulomb != 1) || s->lj_pme || s->vdw_mod != 0)
    {
        return -1;
    }
    s->rvdw      = r;
    real p3      = r * (r * r);
    real p6      = p3 * p3;
    s->cpot_disp = (real)(-1.0 / p6);
    s->cpot_rep  = (real)(-1.0 / (p6 * p6));
    return 0;
}

/* The LJ modifier and its constants (init_interaction_const,
 * forcerec.cpp:848-873; force_switch_constants :787-801,
 * potential_switch_constants :803-818), in double from the real radii and
 * rounded to real once, as the reference's real-valued pow expressions do.
 * Returns -1 for an unknown modifier or a switch radius outside [0, rc). */
int orc_set_vdw_modifier(orc_sys* s, int modifier, double rvdw_switch)
{
    double rc  = (double)s->rvdw;
    double rsw = (double)(real)rvdw_switch;
    double d   = rc - rsw;
    if (modifier < 0 || modifier > 2 || (modifier != 0 && !(rsw >= 0 && rsw < rc)))
    {
        return -1;
    }
    s->vdw_mod     = modifier;
    s->rvdw_switch = (real)rsw;
    s->fsw_c2d = s->fsw_c3d = s->fsw_c2r = s->fsw_c3r = 0;
    s->sw_c3 = s->sw_c4 = s->sw_c5 = 0;
    if (modifier == 0)
    {
        real p3      = s->rvdw * (s->rvdw * s->rvdw);
        real p6      = p3 * p3;
        s->cpot_disp = (real)(-1.0 / p6);
        s->cpot_rep  = (real)(-1.0 / (p6 * p6));
    }
    else if (modifier == 1)
    {
        /* force/p = r^-(p+1) + c2 rsw^2 + c3 rsw^3, potential shifted to 0 at rc */
        double pw[2] = { 6.0, 12.0 };
        real   c2[2], c3[2], cpot[2];
        for (int k = 0; k < 2; k++)
        {
            double p = pw[k];
            c2[k]    = (real)(((p + 1) * rsw - (p + 4) * rc) / (pow(rc, p + 2) * d * d));
            c3[k]    = (real)(-((p + 1) * rsw - (p + 3) * rc) / (pow(rc, p + 2) * d * d * d));
            cpot[k]  = (real)(-pow(rc, -p) + p * (double)c2[k] / 3 * d * d * d + p * (double)c3[k] / 4 * d * d * d * d);
        }
        s->fsw_c2d   = c2[0];
        s->fsw_c3d   = c3[0];
        s->fsw_c2r   = c2[1];
        s->fsw_c3r   = c3[1];
        s->cpot_disp = cpot[0];
        s->cpot_rep  = cpot[1];
    }
    else
    {
        /* sw = 1 + c3 rsw^3 + c4 rsw^4 + c5 rsw^5: 1 at rvdw_switch, 0 at rc */
        s->sw_c3     = (real)(-10 / (d * d * d));
        s->sw_c4     = (real)(15 / (d * d * d * d));
        s->sw_c5     = (real)(-6 / (d * d * d * d * d));
        s->cpot_disp = 0;
        s->cpot_rep  = 0;
    }
    return 0;
}

/* g(x) = exp(-x^2) (1 + x^2 + x^4 / 2) at x = beta rc
 * (compute_lj_function, ewald/ewald_utils.cpp:76-85) */
static real lj_grid_tail(real beta, real rc)
{
    real xrc  = beta * rc;
    real xrc2 = xrc * xrc;
    real xrc4 = xrc2 * xrc2;
    return (real)(exp(-xrc2) * (1 + xrc2 + xrc4 / 2.0));
}

/* calc_ewaldcoeff_lj (ewald_utils.cpp:87-115): doubling, then bisection */
real orc_ewald_coeff_lj(real rc, real rtol)
{
    real beta = 5, low, high;
    int  n, i = 0;
    do
    {
        i++;
        beta *= 2.0;
    } while (lj_grid_tail(beta, rc) > rtol);
    n    = i + 60;
    low  = 0;
    high = beta;
    for (i = 0; i < n; ++i)
    {
        beta = (low + high) / 2.0;
        if (lj_grid_tail(beta, rc) > rtol)
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

/* LJ-PME's real-space part with the potential shift (forcerec.cpp:703-717,
 * kernel_ref_outer.h:182-189).  Returns -1 for an unknown rule or a
 * modifier other than the potential shift. */
int orc_set_lj_pme(orc_sys* s, int comb_rule, double rtol_lj)
{
    if ((comb_rule != 0 && comb_rule != 1) || s->vdw_mod != 0 || !(rtol_lj > 0 && rtol_lj < 1)
        || s->rvdw != s->rcoulomb)
    {
        return -1;
    }
    real rc         = s->rcoulomb;
    real beta       = orc_ewald_coeff_lj(rc, (real)rtol_lj);
    real crc2       = (beta * rc) * (beta * rc);
    real p3         = rc * (rc * rc);
    s->lj_pme       = 1;
    s->ljpme_comb   = comb_rule;
    s->lje_coeff2   = beta * beta;
    s->lje_coeff6_6 = (real)(s->lje_coeff2 * s->lje_coeff2 * s->lje_coeff2 / 6.0);
    s->lje_vc       = (real)((exp(-crc2) * (1 + crc2 + 0.5 * crc2 * crc2) - 1) / (p3 * p3));
    return 0;
}

real orc_lj_pme_beta(const orc_sys* s)
{
    return s->lj_pme ? (real)sqrt((double)s->lje_coeff2) : 0;
}

void orc_destroy(orc_sys* s)
{
    if (!s)
    {
        return;
    }
    free(s->tabF);
    free(s->tabV);
    free(s->tabFDV0);
    free(s->nbfp);
    free(s->excl_off);
    free(s->excl_idx);
    free(s->cxy_na);
    free(s->cxy_ind);
    free(s->atomIndices);
    free(s->cells);
    free(s->xnb);
    free(s->bb);
    free(s->nbtype);
    free(s->nbq);
    free(s->ci);
    free(s->cj);
    free(s->excl);
    free(s->ciO);
    free(s->cjO);
    free(s->exclO);
    free(s->fnb);
    free(s);
}

void orc_get_constants(const orc_sys* s, double* out)
{
    out[0] = s->beta;
    out[1] = s->tab_scale;
    out[2] = s->tab_n;
    out[3] = s->epsfac;
    out[4] = s->k_rf;
    out[5] = s->c_rf;
    out[6] = s->sh_ewald;
    out[7] = s->cpot_disp;
    out[8] = s->cpot_rep;
}

/* rvdw, rvdw_switch, force-switch c2/c3 (dispersion, repulsion),
 * potential-switch c3..c5, LJ-PME beta^2, beta^6/6, sh_lj_ewald */
void orc_get_vdw_constants(const orc_sys* s, double* out)
{
    const real v[12] = { s->rvdw,   s->rvdw_switch, s->fsw_c2d,    s->fsw_c3d,      s->fsw_c2r, s->fsw_c3r,
                         s->sw_c3, s->sw_c4,       s->sw_c5,      s->lje_coeff2,   s->lje_coeff6_6, s->lje_vc };
    for (int k = 0; k < 12; k++)
    {
        out[k] = (double)v[k];
    }
}

void orc_get_tables(const orc_sys* s, real* F, real* V, real* FDV0)
{
    memcpy(F, s->tabF, sizeof(real) * s->tab_n);
    memcpy(V, s->tabV, sizeof(real) * s->tab_n);
    memcpy(FDV0, s->tabFDV0, sizeof(real) * 4 * s->tab_n);
}

/* nbnxn_atomdata_params_init (atomdata.cpp:440-620): copy the ntype^2 table
 * of {6*C6, 12*C12} and append a zero padding type. */
void orc_set_nbfp(orc_sys* s, int ntype, const real* nbfp)
{
    s->numTypes = ntype + 1;
    s->nbfp     = (real*)xrealloc(s->nbfp, sizeof(real) * 2 * s->numTypes * s->numTypes);
    for (int i = 0; i < s->numTypes; i++)
    {
        for (int j = 0; j < s->numTypes; j++)
        {
            int o            = (i * s->numTypes + j) * 2;
            s->nbfp[o]     = (i < ntype && j < ntype) ? nbfp[(i * ntype + j) * 2] : 0;
            s->nbfp[o + 1] = (i < ntype && j < ntype) ? nbfp[(i * ntype + j) * 2 + 1] : 0;
        }
    }
}

void orc_set_exclusions(orc_sys* s, int natoms, const int* off, const int* idx)
{
    s->natoms   = natoms;
    s->excl_off = (int*)xrealloc(s->excl_off, sizeof(int) * (natoms + 1));
    memcpy(s->excl_off, off, sizeof(int) * (natoms + 1));
    s->excl_idx = (int*)xrealloc(s->excl_idx, sizeof(int) * (off[natoms] > 0 ? off[natoms] : 1));
    memcpy(s->excl_idx, idx, sizeof(int) * off[natoms]);
}

/* ------------------------------------------------------------------ */
/* Grid (grid.cpp)                                                     */
/* ------------------------------------------------------------------ */

/* sort_atoms, grid.cpp:292-427: the pigeonhole sort produces the order of
 * ascending (x[dim], atom index).  Restated here with the same pigeonhole
 * insertion so the result is the reference's order by construction. */
static void sort_atoms_z(int* a, int n, const real* x, real h0, real invh, int n_per_h, int* sort)
{
    const int c_sortGridRatio = 4;
    if (n <= 1)
    {
        return;
    }
    invh *= n_per_h * c_sortGridRatio;
    int nsort  = n_per_h * c_sortGridRatio + n;
    for (int i = 0; i < nsort; i++)
    {
        sort[i] = -1;
    }
    for (int i = 0; i < n; i++)
    {
        int zi = (int)((x[a[i] * 3 + 2] - h0) * invh);
        if (zi < 0)
        {
            zi = 0;
        }
        if (zi > n_per_h * c_sortGridRatio)
        {
            zi = n_per_h * c_sortGridRatio;
        }
        if (sort[zi] < 0)
        {
            sort[zi] = a[i];
        }
        else
        {
            while (sort[zi] >= 0
                   && (x[a[i] * 3 + 2] > x[sort[zi] * 3 + 2]
                       || (x[a[i] * 3 + 2] == x[sort[zi] * 3 + 2] && a[i] > sort[zi])))
            {
                zi++;
            }
            if (sort[zi] >= 0)
            {
                int cp  = sort[zi];
                int zim = zi + 1;
                while (sort[zim] >= 0)
                {
                    int tmp   = sort[zim];
                    sort[zim] = cp;
                    cp        = tmp;
                    zim++;
                }
                sort[zim] = cp;
            }
            sort[zi] = a[i];
        }
    }
    int c = 0;
    for (int zi = 0; zi < nsort; zi++)
    {
        if (sort[zi] >= 0)
        {
            a[c++] = sort[zi];
        }
    }
}

/* calc_shifts, pbcutil/pbc.cpp:1187-1202, for a rectangular box */
static void calc_shifts(orc_sys* s)
{
    for (int n = 0, m = -1; m <= 1; m++)
    {
        for (int l = -1; l <= 1; l++)
        {
            for (int k = -2; k <= 2; k++, n++)
            {
                s->shiftvec[n][0] = k * s->box[0];
                s->shiftvec[n][1] = l * s->box[1];
                s->shiftvec[n][2] = m * s->box[2];
            }
        }
    }
}

void orc_get_shiftvec(const orc_sys* s, real* out)
{
    memcpy(out, s->shiftvec, sizeof(s->shiftvec));
}

/* nbnxn_put_on_grid for a rectangular box with lower corner 0 as nblib calls
 * it (api/nblib/gmxcalculator.cpp:85-103, gridset.cpp:135-241,
 * grid.cpp:103-262,1173-1445,988-1048,860-986). */
int orc_put_on_grid(orc_sys* s, const real* box3, int natoms, const real* x)
{
    if (natoms != s->natoms)
    {
        return -1;
    }
    s->box[0] = box3[0];
    s->box[1] = box3[1];
    s->box[2] = box3[2];
    calc_shifts(s);
    /* det(box) in real, then the density nblib passes */
    real det   = s->box[0] * (s->box[1] * s->box[2]);
    s->density = (real)natoms / det;

    /* Grid::setDimensions, simple geometry: 4 atoms per cell */
    if (natoms > ORC_NC)
    {
        real tlen = R_CBRT(ORC_NC / s->density);
        s->ncx    = (int)(s->box[0] / tlen);
        s->ncy    = (int)(s->box[1] / tlen);
        if (s->ncx < 1)
        {
            s->ncx = 1;
        }
        if (s->ncy < 1)
        {
            s->ncy = 1;
        }
    }
    else
    {
        s->ncx = 1;
        s->ncy = 1;
    }
    for (int d = 0; d < 2; d++)
    {
        s->cellSize[d]    = s->box[d] / (d == 0 ? s->ncx : s->ncy);
        s->invCellSize[d] = 1 / s->cellSize[d];
    }
    s->ncol    = s->ncx * s->ncy;
    s->cxy_na  = (int*)xrealloc(s->cxy_na, sizeof(int) * (s->ncol + 1));
    s->cxy_ind = (int*)xrealloc(s->cxy_ind, sizeof(int) * (s->ncol + 2));
    s->cells   = (int*)xrealloc(s->cells, sizeof(int) * (natoms > 0 ? natoms : 1));

    /* calcColumnIndices */
    for (int i = 0; i <= s->ncol; i++)
    {
        s->cxy_na[i] = 0;
    }
    for (int i = 0; i < natoms; i++)
    {
        int cx = (int)((x[i * 3 + 0] - 0) * s->invCellSize[0]);
        int cy = (int)((x[i * 3 + 1] - 0) * s->invCellSize[1]);
        if (cx > s->ncx - 1)
        {
            cx = s->ncx - 1;
        }
        if (cy > s->ncy - 1)
        {
            cy = s->ncy - 1;
        }
        if (cx < 0 || cy < 0)
        {
            return -2; /* the reference requires atoms inside the grid */
        }
        s->cells[i] = cx * s->ncy + cy;
        s->cxy_na[s->cells[i]]++;
    }

    /* setCellIndices */
    int ncz_max = 0;
    s->cxy_ind[0] = 0;
    for (int i = 0; i < s->ncol; i++)
    {
        int ncz = (s->cxy_na[i] + ORC_NC - 1) / ORC_NC;
        if (ncz > ncz_max)
        {
            ncz_max = ncz;
        }
        s->cxy_ind[i + 1] = s->cxy_ind[i] + ncz;
    }
    s->cxy_ind[s->ncol + 1] = s->cxy_ind[s->ncol];
    s->ncells               = s->cxy_ind[s->ncol];
    int nslot               = s->ncells * ORC_NC;
    s->atomIndices          = (int*)xrealloc(s->atomIndices, sizeof(int) * (nslot > 0 ? nslot : 1));
    s->xnb                  = (real*)xrealloc(s->xnb, sizeof(real) * 3 * (nslot > 0 ? nslot : 1));
    s->bb                   = (float*)xrealloc(s->bb, sizeof(float) * 6 * (s->ncells > 0 ? s->ncells : 1));
    s->fnb                  = (real*)xrealloc(s->fnb, sizeof(real) * 3 * (nslot > 0 ? nslot : 1));
    int* fill               = (int*)calloc(s->ncol + 1, sizeof(int));
    for (int i = 0; i < natoms; i++)
    {
        int cxy = s->cells[i];
        s->atomIndices[s->cxy_ind[cxy] * ORC_NC + fill[cxy]++] = i;
    }
    free(fill);
    int  sortsize = (ncz_max * ORC_NC) * 4 + ncz_max * ORC_NC + 8;
    int* sortbuf  = (int*)xrealloc(NULL, sizeof(int) * sortsize);

    /* sortColumnsCpuGeometry + fillCell */
    for (int cxy = 0; cxy < s->ncol; cxy++)
    {
        int na        = s->cxy_na[cxy];
        int ncz       = s->cxy_ind[cxy + 1] - s->cxy_ind[cxy];
        int atomOff   = s->cxy_ind[cxy] * ORC_NC;
        sort_atoms_z(s->atomIndices + atomOff, na, x, 0, (real)(1.0 / (double)s->box[2]), ncz * ORC_NC, sortbuf);
        for (int ind = na; ind < ncz * ORC_NC; ind++)
        {
            s->atomIndices[atomOff + ind] = -1;
        }
        for (int cz = 0; cz < ncz; cz++)
        {
            int c      = s->cxy_ind[cxy] + cz;
            int a0     = c * ORC_NC;
            int nreal  = na - cz * ORC_NC;
            if (nreal > ORC_NC)
            {
                nreal = ORC_NC;
            }
            for (int k = 0; k < ORC_NC; k++)
            {
                int a = s->atomIndices[a0 + k];
                if (k < nreal)
                {
                    s->cells[a] = a0 + k;
                    for (int d = 0; d < 3; d++)
                    {
                        s->xnb[(a0 + k) * 3 + d] = x[a * 3 + d];
                    }
                }
                else
                {
                    for (int d = 0; d < 3; d++)
                    {
                        s->xnb[(a0 + k) * 3 + d] = -1000000;
                    }
                }
            }
            /* calc_bounding_box over the real atoms (grid.cpp:454-484) */
            real lo[3], hi[3];
            for (int d = 0; d < 3; d++)
            {
                lo[d] = hi[d] = s->xnb[a0 * 3 + d];
            }
            for (int k = 1; k < nreal; k++)
            {
                for (int d = 0; d < 3; d++)
                {
                    real v = s->xnb[(a0 + k) * 3 + d];
                    lo[d]  = v < lo[d] ? v : lo[d];
                    hi[d]  = v > hi[d] ? v : hi[d];
                }
            }
            for (int d = 0; d < 3; d++)
            {
#ifdef ORC_DOUBLE
                /* R2F_D / R2F_U, grid.cpp:429-441 */
                s->bb[c * 6 + d]     = (float)(lo[d] >= 0 ? (1 - ORC_FLOAT_EPS) * lo[d] : (1 + ORC_FLOAT_EPS) * lo[d]);
                s->bb[c * 6 + 3 + d] = (float)(hi[d] >= 0 ? (1 + ORC_FLOAT_EPS) * hi[d] : (1 - ORC_FLOAT_EPS) * hi[d]);
#else
                s->bb[c * 6 + d]     = lo[d];
                s->bb[c * 6 + 3 + d] = hi[d];
#endif
            }
        }
    }
    free(sortbuf);
    return 0;
}

/* nbnxn_atomdata_set (atomdata.cpp:955-976): types and charges in grid order,
 * fillers get the padding type and q = 0. */
void orc_set_atom_props(orc_sys* s, const int* type, const real* q)
{
    int nslot = s->ncells * ORC_NC;
    s->nbtype = (int*)xrealloc(s->nbtype, sizeof(int) * (nslot > 0 ? nslot : 1));
    s->nbq    = (real*)xrealloc(s->nbq, sizeof(real) * (nslot > 0 ? nslot : 1));
    for (int k = 0; k < nslot; k++)
    {
        int a        = s->atomIndices[k];
        s->nbtype[k] = a >= 0 ? type[a] : s->numTypes - 1;
        s->nbq[k]    = a >= 0 ? q[a] : 0;
    }
}

/* nbnxn_atomdata_copy_x_to_nbat_x (atomdata.cpp:1018-1073): gather by
 * atomIndices, fillers untouched. */
void orc_convert_coordinates(orc_sys* s, const real* x)
{
    int nslot = s->ncells * ORC_NC;
    for (int k = 0; k < nslot; k++)
    {
        int a = s->atomIndices[k];
        if (a >= 0)
        {
            for (int d = 0; d < 3; d++)
            {
                s->xnb[k * 3 + d] = x[a * 3 + d];
            }
        }
    }
}

int orc_get_grid_dims(const orc_sys* s, int* out)
{
    out[0] = s->ncx;
    out[1] = s->ncy;
    out[2] = s->ncells;
    return 0;
}

void orc_get_grid(const orc_sys* s, int* atomIndices, int* cells, int* cxy_ind, float* bb, real* xnb)
{
    int nslot = s->ncells * ORC_NC;
    if (atomIndices)
    {
        memcpy(atomIndices, s->atomIndices, sizeof(int) * nslot);
    }
    if (cells)
    {
        memcpy(cells, s->cells, sizeof(int) * s->natoms);
    }
    if (cxy_ind)
    {
        memcpy(cxy_ind, s->cxy_ind, sizeof(int) * (s->ncol + 1));
    }
    if (bb)
    {
        memcpy(bb, s->bb, sizeof(float) * 6 * s->ncells);
    }
    if (xnb)
    {
        memcpy(xnb, s->xnb, sizeof(real) * 3 * nslot);
    }
}

/* ------------------------------------------------------------------ */
/* Pair search (pairlist.cpp)                                          */
/* ------------------------------------------------------------------ */

static void add_ci(orc_sys* s, int ci, int shift)
{
    if (s->nci == s->cap_ci)
    {
        s->cap_ci = s->cap_ci ? 2 * s->cap_ci : 1024;
        s->ci     = (int*)xrealloc(s->ci, sizeof(int) * 4 * s->cap_ci);
    }
    int* e = s->ci + 4 * s->nci++;
    e[0]   = ci;
    e[1]   = shift;
    e[2]   = s->ncj;
    e[3]   = s->ncj;
}

static void add_cj(orc_sys* s, int cj, unsigned excl)
{
    if (s->ncj == s->cap_cj)
    {
        s->cap_cj = s->cap_cj ? 2 * s->cap_cj : 4096;
        s->cj     = (int*)xrealloc(s->cj, sizeof(int) * s->cap_cj);
        s->excl   = (unsigned*)xrealloc(s->excl, sizeof(unsigned) * s->cap_cj);
    }
    s->cj[s->ncj]     = cj;
    s->excl[s->ncj++] = excl;
}

/* clusterBoundingBoxDistance2, pairlist.cpp:351-372 */
static float bb_dist2(const float* bi, const float* bj)
{
    float d2 = 0;
    for (int d = 0; d < 3; d++)
    {
        float dl  = bi[d] - bj[3 + d];
        float dh  = bj[d] - bi[3 + d];
        float dm  = dl > dh ? dl : dh;
        float dm0 = dm > 0.0F ? dm : 0.0F;
        d2 += dm0 * dm0;
    }
    return d2;
}

/* The canonical squared distance of an i-atom (already shifted) to a j-atom. */
static real pair_d2(const real* xi, const real* xj)
{
    real dx = xi[0] - xj[0];
    real dy = xi[1] - xj[1];
    real dz = xi[2] - xj[2];
    real a  = dx * dx;
    real b  = dy * dy;
    real c  = dz * dz;
    return (a + b) + c;
}

/* makeClusterListSimple, pairlist.cpp:963-1078 */
static void make_cluster_list_simple(orc_sys* s, const float* bb_ci, const real* x_ci, int icluster,
                                     int jFirst, int jLast, int excludeSubDiagonal, real rlist2, float rbb2)
{
    int inRange = 0;
    while (!inRange && jFirst <= jLast)
    {
        real d2 = bb_dist2(bb_ci, s->bb + 6 * jFirst);
        if (d2 < rbb2)
        {
            inRange = 1;
        }
        else if (d2 < rlist2)
        {
            for (int i = 0; i < ORC_NC && !inRange; i++)
            {
                for (int j = 0; j < ORC_NC; j++)
                {
                    inRange = inRange || pair_d2(x_ci + 3 * i, s->xnb + 3 * (jFirst * ORC_NC + j)) < rlist2;
                }
            }
        }
        if (!inRange)
        {
            jFirst++;
        }
    }
    if (!inRange)
    {
        return;
    }
    inRange = 0;
    while (!inRange && jLast > jFirst)
    {
        real d2 = bb_dist2(bb_ci, s->bb + 6 * jLast);
        if (d2 < rbb2)
        {
            inRange = 1;
        }
        else if (d2 < rlist2)
        {
            for (int i = 0; i < ORC_NC && !inRange; i++)
            {
                for (int j = 0; j < ORC_NC; j++)
                {
                    inRange = inRange || pair_d2(x_ci + 3 * i, s->xnb + 3 * (jLast * ORC_NC + j)) < rlist2;
                }
            }
        }
        if (!inRange)
        {
            jLast--;
        }
    }
    for (int jc = jFirst; jc <= jLast; jc++)
    {
        add_cj(s, jc, (excludeSubDiagonal && icluster == jc) ? ORC_MASK_DIAG : ORC_MASK_ALL);
    }
    s->ci[4 * (s->nci - 1) + 3] = s->ncj;
}

/* get_cell_range, pairlist.cpp:286-309 (maxAtomGroupRadius = 0) */
static void get_cell_range(const orc_sys* s, int dim, real b0, real b1, real d2, real rlist, int* cf, int* cl)
{
    real listRange2       = rlist * rlist;
    real distanceInCells  = (b0 - 0) * s->invCellSize[dim];
    int  nc               = dim == 0 ? s->ncx : s->ncy;
    *cf                   = (int)distanceInCells > 0 ? (int)distanceInCells : 0;
    while (*cf > 0)
    {
        real t = (b0 - 0) - (*cf - 1 + 1) * s->cellSize[dim];
        if (!(d2 + t * t < listRange2))
        {
            break;
        }
        (*cf)--;
    }
    int l = (int)((b1 - 0) * s->invCellSize[dim]);
    *cl   = l < nc - 1 ? l : nc - 1;
    while (*cl < nc - 1)
    {
        real t = (*cl + 1) * s->cellSize[dim] - (b1 - 0);
        if (!(d2 + t * t < listRange2))
        {
            break;
        }
        (*cl)++;
    }
}

/* findJClusterInJList: linear scan of the open i-entry */
static int find_cj(const orc_sys* s, int start, int end, int cj)
{
    for (int k = start; k < end; k++)
    {
        if (s->cj[k] == cj)
        {
            return k;
        }
    }
    return -1;
}

/* setExclusionsForIEntry, pairlist.cpp:1367-1435 */
static void set_exclusions_for_ientry(orc_sys* s, int e, int diagRemoved)
{
    int ci = s->ci[4 * e], start = s->ci[4 * e + 2], end = s->ci[4 * e + 3];
    if (start == end)
    {
        return;
    }
    for (int i = 0; i < ORC_NC; i++)
    {
        int iIndex = ci * ORC_NC + i;
        int iAtom  = s->atomIndices[iIndex];
        if (iAtom < 0)
        {
            continue;
        }
        for (int k = s->excl_off[iAtom]; k < s->excl_off[iAtom + 1]; k++)
        {
            int jAtom = s->excl_idx[k];
            if (jAtom == iAtom)
            {
                continue;
            }
            int jIndex = s->cells[jAtom];
            if (diagRemoved && jIndex <= iIndex)
            {
                continue;
            }
            int jCluster = jIndex / ORC_NC;
            int idx      = find_cj(s, start, end, jCluster);
            if (idx >= 0)
            {
                s->excl[idx] &= ~(1U << (i * ORC_NC + (jIndex - jCluster * ORC_NC)));
            }
        }
    }
}

/* sort_cj_excl / closeIEntry, pairlist.cpp:2002-2067 */
static void close_ientry(orc_sys* s)
{
    int* e     = s->ci + 4 * (s->nci - 1);
    int  start = e[2], end = e[3], n = end - start;
    if (n == 0)
    {
        s->nci--;
        return;
    }
    int       jnew = 0;
    int*      tcj  = (int*)malloc(sizeof(int) * n);
    unsigned* tex  = (unsigned*)malloc(sizeof(unsigned) * n);
    for (int j = start; j < end; j++)
    {
        if (s->excl[j] != ORC_MASK_ALL)
        {
            tcj[jnew]   = s->cj[j];
            tex[jnew++] = s->excl[j];
        }
    }
    if (!((jnew == 0) || (jnew == 1 && s->excl[start] != ORC_MASK_ALL)))
    {
        for (int j = start; j < end; j++)
        {
            if (s->excl[j] == ORC_MASK_ALL)
            {
                tcj[jnew]   = s->cj[j];
                tex[jnew++] = s->excl[j];
            }
        }
        memcpy(s->cj + start, tcj, sizeof(int) * n);
        memcpy(s->excl + start, tex, sizeof(unsigned) * n);
    }
    free(tcj);
    free(tex);
}

/* nbnxn_make_pairlist_part for one intra-grid simple list with full PBC
 * (pairlist.cpp:3076-3582), ci loop in cell order. */
int orc_construct_pairlist(orc_sys* s)
{
    s->nci         = 0;
    s->ncj         = 0;
    s->haveOuter   = 0;
    real  rlist    = s->rlist;
    real  rlist2   = rlist * rlist;
    /* boundingbox_only_distance2, pairlist.cpp:2841-2878 */
    real  bbx      = (real)(0.5 * (s->cellSize[0] + s->cellSize[0]));
    real  bby      = (real)(0.5 * (s->cellSize[1] + s->cellSize[1]));
    real  rbbr     = (real)(rlist - 0.5 * sqrt((double)(bbx * bbx + bby * bby)));
    if (rbbr < 0)
    {
        rbbr = 0;
    }
    real rbb2r = rbbr * rbbr;
#ifdef ORC_DOUBLE
    float rbb2 = (float)((1 + ORC_FLOAT_EPS) * rbb2r);
#else
    float rbb2 = rbb2r;
#endif
    int shp[3] = { 1, 1, 1 };
    if (s->box[0] < rlist)
    {
        shp[0] = 2;
    }
    real listRangeBBToJCell2 = rlist * rlist;

    for (int cix = 0; cix < s->ncx; cix++)
    {
        for (int ciy = 0; ciy < s->ncy; ciy++)
        {
            int ci_xy = cix * s->ncy + ciy;
            for (int ci = s->cxy_ind[ci_xy]; ci < s->cxy_ind[ci_xy + 1]; ci++)
            {
                const float* bbi    = s->bb + 6 * ci;
                int          ncz_i  = s->cxy_ind[ci_xy + 1] - s->cxy_ind[ci_xy];
                for (int tz = -shp[2]; tz <= shp[2]; tz++)
                {
                    real shz = (real)tz * s->box[2];
                    real bz0 = bbi[2] + shz;
                    real bz1 = bbi[5] + shz;
                    real d2z;
                    if (tz == 0)
                    {
                        d2z = 0;
                    }
                    else if (tz < 0)
                    {
                        d2z = bz1 * bz1;
                    }
                    else
                    {
                        d2z = (bz0 - s->box[2]) * (bz0 - s->box[2]);
                    }
                    real d2z_cx = d2z + 0;
                    if (d2z_cx >= rlist2)
                    {
                        continue;
                    }
                    real bz1_frac = bz1 / (real)ncz_i;
                    if (bz1_frac < 0)
                    {
                        bz1_frac = 0;
                    }
                    for (int ty = -shp[1]; ty <= shp[1]; ty++)
                    {
                        real shy = (real)ty * s->box[1] + (real)tz * 0;
                        real by0 = bbi[1] + shy;
                        real by1 = bbi[4] + shy;
                        int  cyf, cyl;
                        get_cell_range(s, 1, by0, by1, d2z_cx, rlist, &cyf, &cyl);
                        if (cyf > cyl)
                        {
                            continue;
                        }
                        for (int tx = -shp[0]; tx <= shp[0]; tx++)
                        {
                            int shift              = xyz2is(tx, ty, tz);
                            int excludeSubDiagonal = (shift == ORC_CENTRAL);
                            if (shift > ORC_CENTRAL)
                            {
                                continue;
                            }
                            real shx = (real)tx * s->box[0] + (real)ty * 0 + (real)tz * 0;
                            real bx0 = bbi[0] + shx;
                            real bx1 = bbi[3] + shx;
                            real d2z_cy = d2z;
                            if (by1 < 0)
                            {
                                d2z_cy += (0 - by1) * (0 - by1);
                            }
                            else if (by0 > s->box[1])
                            {
                                d2z_cy += (by0 - s->box[1]) * (by0 - s->box[1]);
                            }
                            int cxf, cxl;
                            get_cell_range(s, 0, bx0, bx1, d2z_cy, rlist, &cxf, &cxl);
                            if (cxf > cxl)
                            {
                                continue;
                            }
                            add_ci(s, ci, shift);
                            if (excludeSubDiagonal && cxf < cix)
                            {
                                cxf = cix;
                            }
                            /* set_icell_bb_simple / icell_set_x_simple */
                            float bb_ci[6];
                            bb_ci[0] = bbi[0] + shx;
                            bb_ci[1] = bbi[1] + shy;
                            bb_ci[2] = bbi[2] + shz;
                            bb_ci[3] = bbi[3] + shx;
                            bb_ci[4] = bbi[4] + shy;
                            bb_ci[5] = bbi[5] + shz;
                            real x_ci[ORC_NC * 3];
                            for (int i = 0; i < ORC_NC; i++)
                            {
                                x_ci[i * 3 + 0] = s->xnb[(ci * ORC_NC + i) * 3 + 0] + shx;
                                x_ci[i * 3 + 1] = s->xnb[(ci * ORC_NC + i) * 3 + 1] + shy;
                                x_ci[i * 3 + 2] = s->xnb[(ci * ORC_NC + i) * 3 + 2] + shz;
                            }
                            for (int cx = cxf; cx <= cxl; cx++)
                            {
                                real cx_real = (real)cx;
                                real d2zx    = d2z;
                                if (0 + cx_real * s->cellSize[0] > bx1)
                                {
                                    real t = 0 + cx_real * s->cellSize[0] - bx1;
                                    d2zx += t * t;
                                }
                                else if (0 + (cx_real + 1) * s->cellSize[0] < bx0)
                                {
                                    real t = 0 + (cx_real + 1) * s->cellSize[0] - bx0;
                                    d2zx += t * t;
                                }
                                int cyf_x = cyf;
                                if (cx == 0 && shift == ORC_CENTRAL && cyf < ciy)
                                {
                                    cyf_x = ciy;
                                }
                                for (int cy = cyf_x; cy <= cyl; cy++)
                                {
                                    int  col         = cx * s->ncy + cy;
                                    int  columnStart = s->cxy_ind[col];
                                    int  columnEnd   = s->cxy_ind[col + 1];
                                    real cy_real     = (real)cy;
                                    real d2zxy       = d2zx;
                                    if (0 + cy_real * s->cellSize[1] > by1)
                                    {
                                        real t = 0 + cy_real * s->cellSize[1] - by1;
                                        d2zxy += t * t;
                                    }
                                    else if (0 + (cy_real + 1) * s->cellSize[1] < by0)
                                    {
                                        real t = 0 + (cy_real + 1) * s->cellSize[1] - by0;
                                        d2zxy += t * t;
                                    }
                                    if (!(columnStart < columnEnd && d2zxy < listRangeBBToJCell2))
                                    {
                                        continue;
                                    }
                                    int midCell = columnStart + (int)(bz1_frac * (real)(columnEnd - columnStart));
                                    if (midCell >= columnEnd)
                                    {
                                        midCell = columnEnd - 1;
                                    }
                                    real d2xy = d2zxy - d2z;
                                    int  down = midCell;
                                    while (down >= columnStart)
                                    {
                                        real up  = s->bb[6 * down + 5];
                                        real t   = up - bz0;
                                        if (!(up >= bz0 || d2xy + t * t < rlist2))
                                        {
                                            break;
                                        }
                                        down--;
                                    }
                                    int firstCell = down + 1;
                                    int upc       = midCell + 1;
                                    while (upc < columnEnd)
                                    {
                                        real lo = s->bb[6 * upc + 2];
                                        real t  = lo - bz1;
                                        if (!(lo <= bz1 || d2xy + t * t < rlist2))
                                        {
                                            break;
                                        }
                                        upc++;
                                    }
                                    int lastCell = upc - 1;
                                    if (shift == ORC_CENTRAL && firstCell < ci)
                                    {
                                        firstCell = ci;
                                    }
                                    if (firstCell <= lastCell)
                                    {
                                        make_cluster_list_simple(s, bb_ci, x_ci, ci, firstCell, lastCell,
                                                                 excludeSubDiagonal, rlist2, rbb2);
                                    }
                                }
                            }
                            set_exclusions_for_ientry(s, s->nci - 1, excludeSubDiagonal);
                            close_ientry(s);
                        }
                    }
                }
            }
        }
    }
    return s->nci;
}

/* nbnxn_kernel_prune_ref, kernel_ref_prune.cpp:39-123: the list as searched
 * (kept on the first call after a search) pruned to the j-clusters with at
 * least one atom pair, filler atoms included, whose squared distance (i-atom
 * plus shift, minus j-atom, in the canonical rounding) is below
 * rlistInner^2, using the current coordinates; entries keep their order and
 * exclusion masks, i-entries left without j-clusters are dropped. */
int orc_prune_pairlist(orc_sys* s, double rlistInner)
{
    if (!s->haveOuter)
    {
        s->nciO  = s->nci;
        s->ncjO  = s->ncj;
        s->ciO   = (int*)xrealloc(s->ciO, sizeof(int) * 4 * (s->nci + 1));
        s->cjO   = (int*)xrealloc(s->cjO, sizeof(int) * (s->ncj + 1));
        s->exclO = (unsigned*)xrealloc(s->exclO, sizeof(unsigned) * (s->ncj + 1));
        memcpy(s->ciO, s->ci, sizeof(int) * 4 * s->nci);
        memcpy(s->cjO, s->cj, sizeof(int) * s->ncj);
        memcpy(s->exclO, s->excl, sizeof(unsigned) * s->ncj);
        s->haveOuter = 1;
    }
    const real r   = (real)rlistInner;
    const real r2  = r * r;
    s->nci         = 0;
    s->ncj         = 0;
    for (int e = 0; e < s->nciO; e++)
    {
        const int*  ie = s->ciO + 4 * e;
        const int   ci = ie[0], sh = ie[1];
        const real* sv = s->shiftvec[sh];
        real        xi[ORC_NC][3];
        for (int i = 0; i < ORC_NC; i++)
        {
            for (int d = 0; d < 3; d++)
            {
                xi[i][d] = s->xnb[(ci * ORC_NC + i) * 3 + d] + sv[d];
            }
        }
        add_ci(s, ci, sh);
        for (int k = ie[2]; k < ie[3]; k++)
        {
            const int cj      = s->cjO[k];
            int       inRange = 0;
            for (int i = 0; i < ORC_NC && !inRange; i++)
            {
                for (int j = 0; j < ORC_NC; j++)
                {
                    if (pair_d2(xi[i], s->xnb + (cj * ORC_NC + j) * 3) < r2)
                    {
                        inRange = 1;
                    }
                }
            }
            if (inRange)
            {
                add_cj(s, cj, s->exclO[k]);
            }
        }
        int* last = s->ci + 4 * (s->nci - 1);
        last[3]   = s->ncj;
        if (last[3] == last[2])
        {
            s->nci--;
        }
    }
    return s->nci;
}

int orc_num_ci(const orc_sys* s)
{
    return s->nci;
}

int orc_num_cj(const orc_sys* s)
{
    return s->ncj;
}

void orc_get_list(const orc_sys* s, int* ci, int* cj, unsigned* excl)
{
    memcpy(ci, s->ci, sizeof(int) * 4 * s->nci);
    memcpy(cj, s->cj, sizeof(int) * s->ncj);
    memcpy(excl, s->excl, sizeof(unsigned) * s->ncj);
}

/* Expand the cluster list into atom pairs (i-atom, j-atom, shift) whose
 * canonical squared distance is below r2.  Each pair is written as
 * (min atom) << 32 | (max atom).  Returns the number of pairs (writes at
 * most cap). */
int64_t orc_list_pairs(const orc_sys* s, double r2d, int64_t* out, int64_t cap)
{
    real    r2 = (real)r2d;
    int64_t n  = 0;
    for (int e = 0; e < s->nci; e++)
    {
        int        ci = s->ci[4 * e], sh = s->ci[4 * e + 1];
        const real* sv = s->shiftvec[sh];
        for (int k = s->ci[4 * e + 2]; k < s->ci[4 * e + 3]; k++)
        {
            int cj = s->cj[k];
            for (int i = 0; i < ORC_NC; i++)
            {
                int ai = s->atomIndices[ci * ORC_NC + i];
                if (ai < 0)
                {
                    continue;
                }
                real xi[3];
                for (int d = 0; d < 3; d++)
                {
                    xi[d] = s->xnb[(ci * ORC_NC + i) * 3 + d] + sv[d];
                }
                for (int j = 0; j < ORC_NC; j++)
                {
                    int aj = s->atomIndices[cj * ORC_NC + j];
                    if (aj < 0)
                    {
                        continue;
                    }
                    if (sh == ORC_CENTRAL && cj == ci && j <= i)
                    {
                        continue;
                    }
                    if (pair_d2(xi, s->xnb + 3 * (cj * ORC_NC + j)) < r2)
                    {
                        if (n < cap)
                        {
                            int64_t a = ai < aj ? ai : aj, b = ai < aj ? aj : ai;
                            out[n]    = (a << 32) | b;
                        }
                        n++;
                    }
                }
            }
        }
    }
    return n;
}

/* Brute-force O(N^2) enumeration of the same pair set with the same
 * orientation convention as the list: central pairs are orientation-free;
 * shifted pairs are measured from the atom whose shift index is below the
 * central one (pairlist.cpp:3339-3342 keeps only those shifts). */
int64_t orc_brute_pairs(int natoms, const real* x, const real* box3, double r2d, int64_t* out, int64_t cap)
{
    real    r2 = (real)r2d;
    int64_t n  = 0;
    for (int a = 0; a < natoms; a++)
    {
        for (int b = 0; b < natoms; b++)
        {
            if (a == b)
            {
                continue;
            }
            for (int tz = -1; tz <= 1; tz++)
            {
                for (int ty = -1; ty <= 1; ty++)
                {
                    for (int tx = -1; tx <= 1; tx++)
                    {
                        int sh = xyz2is(tx, ty, tz);
                        if (sh > ORC_CENTRAL || (sh == ORC_CENTRAL && b < a))
                        {
                            continue;
                        }
                        real xi[3] = { x[a * 3 + 0] + (real)tx * box3[0], x[a * 3 + 1] + (real)ty * box3[1],
                                       x[a * 3 + 2] + (real)tz * box3[2] };
                        if (pair_d2(xi, x + 3 * b) < r2)
                        {
                            if (n < cap)
                            {
                                int64_t lo = a < b ? a : b, hi = a < b ? b : a;
                                out[n]     = (lo << 32) | hi;
                            }
                            n++;
                        }
                    }
                }
            }
        }
    }
    return n;
}

/* ------------------------------------------------------------------ */
/* Plain-C reference kernel (kernel_ref_outer.h / kernel_ref_inner.h)  */
/* ------------------------------------------------------------------ */

/* The i-entries [e0, e1) of the list: forces into f (cluster layout),
 * shift forces into fshift, energies added to *pVvdw and *pVc. */
static void kernel_entries(orc_sys* s, int energy, const real* ljc, int e0, int e1, real* f, real* fshift,
                           real* pVvdw, real* pVc)
{
    const real rcut2    = s->rcoulomb * s->rcoulomb;
    const int  twin     = s->rvdw < s->rcoulomb;
    const real rvdw2    = s->rvdw * s->rvdw;
    const int  ntype2   = s->numTypes * 2;
    const real facel    = s->epsfac;
    const real k_rf2    = 2 * s->k_rf;
    const real halfsp   = (real)(0.5 / s->tab_scale);
    const int  tab      = (s->coulomb == 1 && !s->ewald_exact);
    const int  exact    = (s->coulomb == 1 && s->ewald_exact);
    const real* x = s->xnb;
    const real* q = s->nbq;
    for (int e = e0; e < e1; e++)
    {
        int  ci    = s->ci[4 * e];
        int  ish   = s->ci[4 * e + 1];
        int  cj0   = s->ci[4 * e + 2];
        int  cj1   = s->ci[4 * e + 3];
        int  ci_sh = (ish == ORC_CENTRAL ? ci : -1);
        real xi[ORC_NC * 3], fi[ORC_NC * 3], qi[ORC_NC];
        real Vvdw_ci = 0, Vc_ci = 0;
        for (int i = 0; i < ORC_NC; i++)
        {
            for (int d = 0; d < 3; d++)
            {
                xi[i * 3 + d] = x[(ci * ORC_NC + i) * 3 + d] + s->shiftvec[ish][d];
                fi[i * 3 + d] = 0;
            }
            qi[i] = facel * q[ci * ORC_NC + i];
        }
        if (energy)
        {
            real Vc_sub_self;
            if (s->coulomb == 0)
            {
                Vc_sub_self = (real)(0.5 * s->c_rf);
            }
            else if (tab)
            {
#ifdef ORC_DOUBLE
                Vc_sub_self = (real)(0.5 * s->tabV[0]);
#else
                Vc_sub_self = (real)(0.5 * s->tabFDV0[2]);
#endif
            }
            else
            {
                Vc_sub_self = (real)(0.5 * orc_ewald_v_exact(s->beta, 0));
            }
            if (s->cj[cj0] == ci_sh)
            {
                for (int i = 0; i < ORC_NC; i++)
                {
                    (*pVc) -= qi[i] * q[ci * ORC_NC + i] * Vc_sub_self;
                    if (s->lj_pme)
                    {
                        /* LJ Ewald self interaction (kernel_ref_outer.h:315-322) */
                        int t = s->nbtype[ci * ORC_NC + i];
                        (*pVvdw) += (real)0.5 * s->nbfp[(t * s->numTypes + t) * 2] / 6 * s->lje_coeff6_6;
                    }
                }
            }
        }
        for (int cjind = cj0; cjind < cj1; cjind++)
        {
            int      cj   = s->cj[cjind];
            unsigned excl = s->excl[cjind];
            for (int i = 0; i < ORC_NC; i++)
            {
                int ai         = ci * ORC_NC + i;
                int type_i_off = s->nbtype[ai] * ntype2;
                for (int j = 0; j < ORC_NC; j++)
                {
                    int  aj       = cj * ORC_NC + j;
                    real interact = (real)((excl >> (i * ORC_NC + j)) & 1);
                    real skipmask = (cj == ci_sh && j <= i) ? 0.0 : 1.0;
                    real dx       = xi[i * 3 + 0] - x[aj * 3 + 0];
                    real dy       = xi[i * 3 + 1] - x[aj * 3 + 1];
                    real dz       = xi[i * 3 + 2] - x[aj * 3 + 2];
                    real rsq      = dx * dx + dy * dy + dz * dz;
                    skipmask      = (rsq >= rcut2) ? 0 : skipmask;
                    rsq           = rsq > ORC_MIN_RSQ ? rsq : ORC_MIN_RSQ;
                    real rinv     = 1 / R_SQRT(rsq);
                    rinv          = rinv * skipmask;
                    real rinvsq   = rinv * rinv;
                    real c6       = s->nbfp[type_i_off + s->nbtype[aj] * 2];
                    real c12      = s->nbfp[type_i_off + s->nbtype[aj] * 2 + 1];
                    real rinvsix  = interact * rinvsq * rinvsq * rinvsq;
                    real FrLJ6    = c6 * rinvsix;
                    real FrLJ12   = c12 * rinvsix * rinvsix;
                    real frLJ     = FrLJ12 - FrLJ6;
                    real VLJ      = 0;
                    real r = 0, rsw = 0;
                    if (energy || s->vdw_mod == 2)
                    {
                        VLJ = (FrLJ12 + c12 * s->cpot_rep) / 12 - (FrLJ6 + c6 * s->cpot_disp) / 6;
                    }
                    if (s->vdw_mod != 0)
                    {
                        /* switching from rvdw_switch (kernel_ref_inner.h:168-172):
                         * r from the masked rinv */
                        r   = rsq * rinv;
                        rsw = r - s->rvdw_switch;
                        rsw = (rsw >= 0 ? rsw : 0);
                    }
                    if (s->vdw_mod == 1)
                    {
                        /* kernel_ref_inner.h:174-183 */
                        frLJ += -c6 * (s->fsw_c2d + s->fsw_c3d * rsw) * rsw * rsw * r
                                + c12 * (s->fsw_c2r + s->fsw_c3r * rsw) * rsw * rsw * r;
                        if (energy)
                        {
                            VLJ += -c6 * (-s->fsw_c2d / 3 - s->fsw_c3d / 4 * rsw) * rsw * rsw * rsw
                                   + c12 * (-s->fsw_c2r / 3 - s->fsw_c3r / 4 * rsw) * rsw * rsw * rsw;
                        }
                    }
                    /* masked after the force switch, before the potential
                     * switch (kernel_ref_inner.h:185-191) */
                    VLJ = VLJ * interact;
                    if (s->vdw_mod == 2)
                    {
                        /* kernel_ref_inner.h:193-203, kernel_ref_outer.h:171-176 */
                        real sw  = 1 + (s->sw_c3 + (s->sw_c4 + s->sw_c5 * rsw) * rsw) * rsw * rsw * rsw;
                        real dsw = (3 * s->sw_c3 + (4 * s->sw_c4 + 5 * s->sw_c5 * rsw) * rsw) * rsw * rsw;
                        frLJ     = frLJ * sw - r * VLJ * dsw;
                        VLJ *= sw;
                    }
                    if (s->lj_pme)
                    {
                        /* kernel_ref_inner.h:207-248 */
                        int  ti = s->nbtype[ai], tj = s->nbtype[aj];
                        real c6grid;
                        if (s->ljpme_comb == 0)
                        {
                            c6grid = ljc[ti * 2] * ljc[tj * 2];
                        }
                        else
                        {
                            real sigma   = ljc[ti * 2] + ljc[tj * 2];
                            real epsilon = ljc[ti * 2 + 1] * ljc[tj * 2 + 1];
                            real sigma2  = sigma * sigma;
                            c6grid       = epsilon * sigma2 * sigma2 * sigma2;
                        }
                        /* r^-6 without the exclusion mask */
                        real rinvsix_nm = rinvsq * rinvsq * rinvsq;
                        real cr2        = s->lje_coeff2 * rsq;
                        real expmcr2    = (real)exp(-(double)cr2);
                        real poly       = 1 + cr2 + (real)0.5 * cr2 * cr2;
                        frLJ += c6grid * (rinvsix_nm - expmcr2 * (rinvsix_nm * poly + s->lje_coeff6_6));
                        if (energy)
                        {
                            real sh_mask = s->lje_vc * interact;
                            VLJ += c6grid / 6 * (rinvsix_nm * (1 - expmcr2 * poly) + sh_mask);
                        }
                    }
                    if (twin)
                    {
                        /* LJ only within rvdw (VDW_CUTOFF_CHECK,
                         * kernel_ref_inner.h:254-265), on the floored r^2 */
                        real skipmask_rvdw = (rsq < rvdw2) ? 1 : 0;
                        frLJ *= skipmask_rvdw;
                        VLJ *= skipmask_rvdw;
                    }
                    else
                    {
                        VLJ = VLJ * skipmask;
                    }
                    if (energy)
                    {
                        Vvdw_ci += VLJ;
                    }
                    real qq = skipmask * qi[i] * q[aj];
                    real fcoul, vcoul = 0;
                    if (s->coulomb == 0)
                    {
                        fcoul = qq * (interact * rinv * rinvsq - k_rf2);
                        if (energy)
                        {
                            vcoul = qq * (interact * rinv + s->k_rf * rsq - s->c_rf);
                        }
                    }
                    else if (tab)
                    {
                        real rs   = rsq * rinv * s->tab_scale;
                        int  ri   = (int)rs;
                        real frac = rs - (real)ri;
#ifndef ORC_DOUBLE
                        real fexcl = s->tabFDV0[ri * 4] + frac * s->tabFDV0[ri * 4 + 1];
#else
                        real fexcl = (1 - frac) * s->tabF[ri] + frac * s->tabF[ri + 1];
#endif
                        fcoul = interact * rinvsq - fexcl;
                        if (energy)
                        {
#ifndef ORC_DOUBLE
                            vcoul = qq * (interact * (rinv - s->sh_ewald)
                                          - (s->tabFDV0[ri * 4 + 2] - halfsp * frac * (s->tabFDV0[ri * 4] + fexcl)));
#else
                            vcoul = qq * (interact * (rinv - s->sh_ewald)
                                          - (s->tabV[ri] - halfsp * frac * (s->tabF[ri] + fexcl)));
#endif
                        }
                        fcoul *= qq * rinv;
                    }
                    else
                    {
                        (void)exact;
                        real r     = rsq * rinv;
                        real fexcl = (real)orc_ewald_f_exact(s->beta, r);
                        fcoul      = interact * rinvsq - fexcl;
                        if (energy)
                        {
                            vcoul = qq * (interact * (rinv - s->sh_ewald) - (real)orc_ewald_v_exact(s->beta, r));
                        }
                        fcoul *= qq * rinv;
                    }
                    if (energy)
                    {
                        Vc_ci += vcoul;
                    }
                    real fscal = frLJ * rinvsq + fcoul;
                    real fx    = fscal * dx;
                    real fy    = fscal * dy;
                    real fz    = fscal * dz;
                    fi[i * 3 + 0] += fx;
                    fi[i * 3 + 1] += fy;
                    fi[i * 3 + 2] += fz;
                    f[aj * 3 + 0] -= fx;
                    f[aj * 3 + 1] -= fy;
                    f[aj * 3 + 2] -= fz;
                }
            }
        }
        for (int i = 0; i < ORC_NC; i++)
        {
            for (int d = 0; d < 3; d++)
            {
                f[(ci * ORC_NC + i) * 3 + d] += fi[i * 3 + d];
            }
        }
        for (int i = 0; i < ORC_NC; i++)
        {
            for (int d = 0; d < 3; d++)
            {
                fshift[ish * 3 + d] += fi[i * 3 + d];
            }
        }
        if (energy)
        {
            (*pVvdw) += Vvdw_ci;
            (*pVc) += Vc_ci;
        }
    }
}

/* LJ-PME grid parameters per type from the nbfp diagonal (nbfp_comb,
 * atomdata.cpp:288-330); the caller frees the array. */
static real* lj_pme_grid_params(const orc_sys* s)
{
    real* ljc = (real*)calloc((size_t)(2 * s->numTypes + 2), sizeof(real));
    for (int t = 0; s->lj_pme && t < s->numTypes; t++)
    {
        real c6  = s->nbfp[(t * s->numTypes + t) * 2];
        real c12 = s->nbfp[(t * s->numTypes + t) * 2 + 1];
        if (s->ljpme_comb == 0)
        {
            ljc[t * 2]     = (real)sqrt((double)c6);
            ljc[t * 2 + 1] = (real)sqrt((double)c12);
        }
        else if (c6 > 0 && c12 > 0)
        {
            ljc[t * 2]     = (real)(0.5 * pow((double)(c12 / c6), 1.0 / 6.0));
            ljc[t * 2 + 1] = (real)sqrt((double)(c6 * c6 / c12));
        }
    }
    return ljc;
}

/* One nbnxn_kernel_Elec{RF,QSTab}_VdwLJ_{F,VF}_ref call over the whole list
 * (a single list, i.e. one OpenMP thread), clearing nbat f first as
 * enbvClearFYes does.  The inner loop is the "LJ + Coulomb with exclusion
 * check" variant (kernel_ref_inner.h with CHECK_EXCLS and CALC_COULOMB): the
 * reference runs every entry through it because its all-ones mask is
 * 0xffffffff while the loop tests for 0xffff (kernel_ref_outer.h:329). */
void orc_kernel(orc_sys* s, int energy)
{
    const int nslot = s->ncells * ORC_NC;
    real*     ljc   = lj_pme_grid_params(s);
    for (int k = 0; k < nslot * 3; k++)
    {
        s->fnb[k] = 0;
    }
    for (int k = 0; k < ORC_SHIFTS * 3; k++)
    {
        s->fshift[k] = 0;
    }
    s->Vvdw = 0;
    s->Vc   = 0;
    kernel_entries(s, energy, ljc, 0, s->nci, s->fnb, s->fshift, &s->Vvdw, &s->Vc);
    free(ljc);
}

#ifdef ORC_OMP
void orc_add_nbat_f_to_f(const orc_sys* s, real* fout);

/* The CPU baseline only (built apart with OpenMP, never the checker): the
 * kernel over the list with OpenMP threads, each with its own force, shift
 * force and energy outputs (nbnxn_atomdata_output_t per thread,
 * kerneldispatch.cpp:243-356), reduced over the threads afterwards as
 * reduceForces does (atomdata.cpp:1425-1469). */
void orc_kernel_threads(orc_sys* s, int energy, int nthreads)
{
    const int nslot = s->ncells * ORC_NC;
    real*     ljc   = lj_pme_grid_params(s);
    real**    fth   = (real**)calloc((size_t)nthreads, sizeof(real*));
    real*     fsh   = (real*)calloc((size_t)nthreads * ORC_SHIFTS * 3, sizeof(real));
    real*     vv    = (real*)calloc((size_t)nthreads * 2, sizeof(real));
#    pragma omp parallel num_threads(nthreads)
    {
        const int th = omp_get_thread_num();
        fth[th]      = (real*)calloc((size_t)nslot * 3, sizeof(real));
#    pragma omp for schedule(dynamic, 32)
        for (int e = 0; e < s->nci; e++)
        {
            kernel_entries(s, energy, ljc, e, e + 1, fth[th], fsh + th * ORC_SHIFTS * 3, vv + 2 * th,
                           vv + 2 * th + 1);
        }
#    pragma omp for schedule(static)
        for (int k = 0; k < nslot * 3; k++)
        {
            real sum = 0;
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
    s->Vvdw = 0;
    s->Vc   = 0;
    for (int t = 0; t < nthreads; t++)
    {
        s->Vvdw += vv[2 * t];
        s->Vc += vv[2 * t + 1];
        free(fth[t]);
    }
    free(fth);
    free(fsh);
    free(vv);
    free(ljc);
}

/* convertCoordinates, the threaded kernel and add_nbat_f_to_f. */
void orc_compute_threads(orc_sys* s, const real* x, real* f, int energy, int nthreads, double* out_e)
{
    orc_convert_coordinates(s, x);
    orc_kernel_threads(s, energy, nthreads);
    orc_add_nbat_f_to_f(s, f);
    out_e[0] = s->Vvdw;
    out_e[1] = s->Vc;
}

/* the AVX-512 baseline kernel (the reference's SIMD scheme) */
#    include "baseline_avx512.c"
#endif

/* add_nbat_f_to_f for one output buffer (atomdata.cpp:1181-1228), with the
 * zeroing nblib does first (gmxcalculator.cpp:76-82): f[a] = fnb[cells[a]]. */
void orc_add_nbat_f_to_f(const orc_sys* s, real* fout)
{
    for (int a = 0; a < s->natoms; a++)
    {
        int c = s->cells[a];
        for (int d = 0; d < 3; d++)
        {
            fout[a * 3 + d] = 0;
            fout[a * 3 + d] += s->fnb[c * 3 + d];
        }
    }
}

/* GmxForceCalculator::compute (gmxcalculator.cpp:70-83) plus the energies,
 * shift forces and virial nblib keeps internal.  out_e = {Vvdw, Vc};
 * vir = -1/2 sum_i x_i (x) f_i - 1/2 sum_s shift_s (x) fshift_s
 * (calcvir.cpp:103-148, sim_util.cpp:157-177), accumulated in double. */
void orc_compute(orc_sys* s, const real* x, real* f, int energy, double* out_e, real* fshift, double* vir)
{
    orc_convert_coordinates(s, x);
    orc_kernel(s, energy);
    orc_add_nbat_f_to_f(s, f);
    if (out_e)
    {
        out_e[0] = s->Vvdw;
        out_e[1] = s->Vc;
    }
    if (fshift)
    {
        memcpy(fshift, s->fshift, sizeof(s->fshift));
    }
    if (vir)
    {
        for (int k = 0; k < 9; k++)
        {
            vir[k] = 0;
        }
        for (int a = 0; a < s->natoms; a++)
        {
            for (int d1 = 0; d1 < 3; d1++)
            {
                for (int d2 = 0; d2 < 3; d2++)
                {
                    vir[d1 * 3 + d2] -= 0.5 * (double)x[a * 3 + d1] * (double)f[a * 3 + d2];
                }
            }
        }
        for (int sh = 0; sh < ORC_SHIFTS; sh++)
        {
            for (int d1 = 0; d1 < 3; d1++)
            {
                for (int d2 = 0; d2 < 3; d2++)
                {
                    vir[d1 * 3 + d2] -= 0.5 * (double)s->shiftvec[sh][d1] * (double)s->fshift[sh * 3 + d2];
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Brute-force FP64 pair sum (independent of grid and list)            */
/* ------------------------------------------------------------------ */

/* Direct sum over all atom pairs with the minimum image, in double, with
 * the same physics conventions: potential-shifted LJ cut, Ewald real space
 * with the exact closed form (or RF), the Ewald/RF correction on excluded
 * pairs inside the cut-off, and the self term -1/2 q_i^2 epsfac V(0).
 * nbfp holds 6*C6 and 12*C12 for ntype x ntype.  Returns energies in
 * out_e = {Vvdw, Vc}; f is natoms*3 (double); vir is the pair virial
 * -1/2 sum r_ij (x) f_ij. */
void orc_brute_forces(int natoms, const double* x, const double* box3, const int* type, const double* q,
                      int ntype, const double* nbfp, const int* excl_off, const int* excl_idx, int coulomb,
                      double rc, double beta, double epsfac, double k_rf, double c_rf, double sh_ewald,
                      double* f, double* out_e, double* vir)
{
    double rc2   = rc * rc;
    double cpd   = -1.0 / pow(rc, 6);
    double cpr   = -1.0 / pow(rc, 12);
    double Vvdw  = 0, Vc = 0;
    for (int k = 0; k < natoms * 3; k++)
    {
        f[k] = 0;
    }
    for (int k = 0; k < 9; k++)
    {
        vir[k] = 0;
    }
    for (int a = 0; a < natoms; a++)
    {
        double vself = coulomb == 1 ? orc_ewald_v_exact(beta, 0) : c_rf;
        Vc -= 0.5 * epsfac * q[a] * q[a] * vself;
        for (int b = a + 1; b < natoms; b++)
        {
            double d[3];
            for (int k = 0; k < 3; k++)
            {
                d[k] = x[a * 3 + k] - x[b * 3 + k];
                d[k] -= box3[k] * floor(d[k] / box3[k] + 0.5);
            }
            double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (rsq >= rc2)
            {
                continue;
            }
            int excluded = 0;
            for (int k = excl_off[a]; k < excl_off[a + 1]; k++)
            {
                if (excl_idx[k] == b)
                {
                    excluded = 1;
                }
            }
            double r     = sqrt(rsq);
            double rinv  = 1 / r;
            double qq    = epsfac * q[a] * q[b];
            double fscal = 0;
            if (!excluded)
            {
                double c6   = nbfp[(type[a] * ntype + type[b]) * 2];
                double c12  = nbfp[(type[a] * ntype + type[b]) * 2 + 1];
                double rinv6 = rinv * rinv * rinv * rinv * rinv * rinv;
                fscal += (c12 * rinv6 * rinv6 - c6 * rinv6) * rinv * rinv;
                Vvdw += c12 * (rinv6 * rinv6 + cpr) / 12 - c6 * (rinv6 + cpd) / 6;
            }
            if (coulomb == 1)
            {
                double fe = orc_ewald_f_exact(beta, r);
                double ve = orc_ewald_v_exact(beta, r);
                if (!excluded)
                {
                    fscal += qq * (rinv * rinv * rinv - fe * rinv);
                    Vc += qq * (rinv - sh_ewald - ve);
                }
                else
                {
                    fscal += qq * (-fe * rinv);
                    Vc += qq * (-ve);
                }
            }
            else
            {
                if (!excluded)
                {
                    fscal += qq * (rinv * rinv * rinv - 2 * k_rf);
                    Vc += qq * (rinv + k_rf * rsq - c_rf);
                }
                else
                {
                    fscal += qq * (-2 * k_rf);
                    Vc += qq * (k_rf * rsq - c_rf);
                }
            }
            for (int k = 0; k < 3; k++)
            {
                f[a * 3 + k] += fscal * d[k];
                f[b * 3 + k] -= fscal * d[k];
                for (int l = 0; l < 3; l++)
                {
                    vir[k * 3 + l] -= 0.5 * d[k] * fscal * d[l];
                }
            }
        }
    }
    out_e[0] = Vvdw;
    out_e[1] = Vc;
}
