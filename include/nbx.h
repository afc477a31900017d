/*
 * nbx.h - C ABI of the MI355X-native NBNxM short-range nonbonded engine.
 *
 * One nbx_context holds one periodic rectangular system on one HIP device:
 * the LJ parameters, the topology (types, charges, exclusions), the grid of
 * 4-atom clusters, the cluster-pair list and the force/energy buffers.  It
 * replaces the reference's nonbonded_verlet_t behind nblib's
 * ForceCalculator (api/nblib/gmxcalculator.h:92-111); the entry points map
 * one to one onto the calls nblib makes (gmxsetup.cpp:306-325 for set-up,
 * gmxcalculator.cpp:70-83 for a step).
 *
 * Plain pointers and sizes only.  Coordinates and forces are float xyz
 * triplets in the caller's atom order; a pointer is a host pointer unless
 * the *_on_device argument says it is device memory of the context's device.
 * Every function returns NBX_SUCCESS or an error code and records a message
 * for nbx_last_error() (per thread).  Calls on one context are not
 * re-entrant; different contexts are independent.
 *
 * Units: nm, kJ/mol, e.  Precision: FP32 pair arithmetic, FP64 energy,
 * shift-force and virial accumulation.
 */
#ifndef NBX_H
#define NBX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBX_SUCCESS 0
/* invalid user input: nblib throws nblib::InputException for these
 * (api/nblib/exception.h; e.g. gmxsetup.cpp:80-98, gmxcalculator.cpp:91-94) */
#define NBX_ERROR_INPUT 1
/* a HIP runtime error */
#define NBX_ERROR_DEVICE 2
/* calls out of order (e.g. compute before a pair list exists) */
#define NBX_ERROR_STATE 3

/* Coulomb flavour: nblib CoulombType {Pme, Cutoff, ReactionField}
 * (api/nblib/kerneloptions.h:74-80).  Cutoff runs the reaction-field kernel
 * with eps_rf = eps_r (forcerec.cpp:902-915, kerneldispatch.cpp:166-169). */
#define NBX_COULOMB_EWALD 0
#define NBX_COULOMB_CUTOFF 1
#define NBX_COULOMB_REACTION_FIELD 2

/* Ewald exclusion correction: the reference's quadratic-spline table
 * (Nbnxm::EwaldExclusionType::Table, what the plain-C kernel always uses,
 * gmxsetup.cpp:117-121) or an analytical evaluation. */
#define NBX_EWALD_TABLE 0
#define NBX_EWALD_ANALYTICAL 1

/* LJ modifier (interaction_const_t::vdw_modifier; constants as
 * mdlib/forcerec.cpp:780-818,858-873 sets them; kernel terms as the
 * plain-C kernel's LJ_FORCE_SWITCH / LJ_POT_SWITCH paths,
 * nbnxm/kernels_reference/kernel_ref_inner.h:155-205).  nblib always uses
 * the potential shift (gmxsetup.cpp:264-267); the switches act between
 * rvdw_switch and the cut-off. */
#define NBX_VDW_POTENTIAL_SHIFT 0
#define NBX_VDW_FORCE_SWITCH 1
#define NBX_VDW_POTENTIAL_SWITCH 2

/* LJ type (interaction_const_t::vdwtype): a plain cut-off, or LJ-PME, whose
 * real-space kernel subtracts the grid's dispersion part with the
 * potential shift (kernel_ref_inner.h:207-248, the self term
 * kernel_ref_outer.h:315-322, ewaldcoeff_lj and sh_lj_ewald
 * forcerec.cpp:703-717).  The grid's C6 per type pair comes from the C6/C12
 * diagonal by the geometric or the Lorentz-Berthelot rule (nbfp_comb,
 * atomdata.cpp:288-330; the rule as nbnxm_setup.cpp:398-410 picks it).
 * The mesh part itself is not computed here (as in nbnxm). */
#define NBX_VDW_CUTOFF 0
#define NBX_VDW_PME 1
#define NBX_LJPME_GEOMETRIC 0
#define NBX_LJPME_LORENTZ_BERTHELOT 1

/* compute flags (gmx::StepWorkload computeForces / computeEnergy /
 * computeVirial, gmxsetup.cpp:214-224) */
#define NBX_COMPUTE_FORCES 1u
#define NBX_COMPUTE_ENERGY 2u
#define NBX_COMPUTE_SHIFT_FORCES 4u
#define NBX_COMPUTE_VIRIAL 8u
/* forces added to f instead of overwriting it: atomdata_add_nbat_f_to_f
 * without nblib's zero fill (nbnxm.cpp:162-180), and on device buffers
 * atomdata_add_nbat_f_to_f_gpu (nbnxm.h:345-359), for a caller that sums
 * several force sources in one buffer */
#define NBX_COMPUTE_ADD_FORCES 16u
/* bitwise reproducible forces, virial and shift forces in every case: the
 * forces of j-clusters that do not fit a group's LDS window (a small share
 * of the cluster pairs at long cut-offs or with large groups) are summed in
 * 64-bit fixed point (2^-20 kJ mol^-1 nm^-1, range +-8.8e12) instead of
 * float atomics, whose order varies from run to run.  No counterpart in the
 * reference's compute flags (its CPU kernels are reproducible and its GPU
 * kernels are not); costs a few per cent of the pair kernel when windows
 * overflow, nothing otherwise. */
#define NBX_COMPUTE_REPRODUCIBLE 32u

typedef struct nbx_context nbx_context;

/* NBKernelOptions subset (api/nblib/kerneloptions.h:86-112) plus the
 * extensions the benchmark needs: a Coulomb cut-off below the list cut-off,
 * the Ewald tolerance and mdrun's potential-shift convention. */
typedef struct nbx_options
{
    int    coulomb_type;          /* NBX_COULOMB_*, default EWALD (nblib Pme) */
    int    ewald_correction;      /* NBX_EWALD_*, default TABLE */
    double pairlist_cutoff;       /* rlist, nm (nblib pairlistCutoff, 1.0) */
    double coulomb_cutoff;        /* rcoulomb <= rlist; 0 means rlist */
    double ewald_rtol;            /* erfc(beta rc) = rtol, 1e-5 (gmxsetup.cpp:275) */
    double epsilon_r;             /* 1 (gmxsetup.cpp:251-252) */
    double epsilon_rf;            /* 1 */
    int    ewald_potential_shift; /* 0: sh_ewald = 0 as nblib; 1: mdrun's shift */
    int    device;                /* HIP device ordinal */
    int    kernel_group_size;     /* i-clusters per kernel work group: 0 = largest
                                     that fits LDS, else 1, 2, 4, 8 or 16 (tuning) */
    int    kernel_blocks_per_cu;  /* resident kernel workgroups per CU: 0 = as many
                                     as LDS allows up to 4 (tuning) */
    int    vdw_modifier;          /* NBX_VDW_*, default POTENTIAL_SHIFT (nblib) */
    double rvdw_switch;           /* switch radius for the switches, nm, 0 <= rvdw_switch < rcoulomb */
    int    vdw_type;              /* NBX_VDW_CUTOFF (nblib) or NBX_VDW_PME (potential shift only) */
    int    ljpme_comb_rule;       /* NBX_LJPME_*: the grid's combination rule, default GEOMETRIC */
    double ewald_rtol_lj;         /* LJ-PME: g(beta_lj rc) = rtol, default 1e-3 (ewald-rtol-lj) */
    double vdw_cutoff;            /* rvdw <= rcoulomb; 0 means rcoulomb.  rvdw < rcoulomb is the
                                     twin cut-off (the coulktTAB_TWIN / coulktEWALD_TWIN kernels,
                                     kerneldispatch.cpp:175-192, VDW_CUTOFF_CHECK in
                                     kernel_ref_inner.h:254-271): Ewald Coulomb with an LJ cut-off
                                     only, as grompp allows (readir.cpp:386-398); the LJ
                                     potential shift and switches then act at rvdw */
} nbx_options;

typedef struct nbx_pairlist_info
{
    int64_t num_atoms;
    int64_t num_clusters;     /* 4-atom clusters incl. partly filled ones */
    int64_t num_columns_x;    /* grid columns (Grid::setDimensions) */
    int64_t num_columns_y;
    int64_t num_i_entries;    /* (i-cluster, shift) entries */
    int64_t num_j_entries;    /* cluster pairs; 16 atom pairs each */
} nbx_pairlist_info;

/* Message of the last failed call on this thread ("" if none). */
const char* nbx_last_error(void);

/* nblib's defaults (NBKernelOptions, kerneloptions.h:86-112). */
void nbx_default_options(nbx_options* options);

/* GmxSetupDirector/NbvSetupUtil (gmxsetup.cpp:100-325): validates options,
 * computes the interaction constants (setupInteractionConst :226-284: Ewald
 * coefficient calc_ewaldcoeff_q ewald_utils.cpp:46-74, correction tables
 * init_interaction_const_tables forcerec.cpp:726-778, LJ potential shifts,
 * reaction-field constants) and binds the device. */
int nbx_create(const nbx_options* options, nbx_context** context);

/* Releases the context and its device memory. */
int nbx_destroy(nbx_context* context);

/* Work is queued on this HIP stream (hipStream_t) from now on; NULL is
 * HIP's default (null) stream.  A new context uses a non-blocking stream of
 * its own.  No counterpart in the reference (CPU only there). */
int nbx_set_stream(nbx_context* context, void* stream);

/* NbvSetupUtil::setNonBondedParameters (gmxsetup.cpp:143-166) and
 * nbnxn_atomdata_init's parameter table (atomdata.cpp:440-620): c6 and c12
 * are num_types x num_types matrices (row-major, unscaled C6/C12); the
 * kernel stores {6*C6, 12*C12} plus a zero padding type for fillers.  New
 * values for the same number of types keep the grid and pair list. */
int nbx_set_nonbonded_parameters(nbx_context* context, int num_types, const double* c6, const double* c12);

/* The Topology inputs nblib feeds to nbnxm (topology.h:72-121): per-atom
 * type ids and charges, and the exclusions as a list of lists in CSR form
 * (gmx::ListOfLists<int>, every atom excluding itself, molecules.cpp:86-88).
 * All arrays are host memory. */
int nbx_set_topology(nbx_context* context, int num_atoms, const int* types, const float* charges,
                     const int* exclusion_offsets, const int* exclusion_indices);

/* nbnxn_put_on_grid (nbnxm.h:453-464) as GmxForceCalculator::setParticlesOnGrid
 * calls it (gmxcalculator.cpp:85-103: lower corner 0, upper corner the box
 * diagonal, density N/det(box)), followed by setAtomProperties
 * (nbnxm.cpp:122-127, atomdata.cpp:955-976).  box holds the three edge
 * lengths of a rectangular box; atoms must lie in [0, L). */
int nbx_put_on_grid(nbx_context* context, const float* box, const float* x, int x_on_device);

/* nonbonded_verlet_t::constructPairlist (pairlist.cpp:4247-4262) for the
 * local intra-grid list with full periodicity. */
int nbx_construct_pairlist(nbx_context* context);

/* nonbonded_verlet_t::dispatchPruneKernelCpu (nbnxm.h:312-313) with the
 * plain-C prune kernel nbnxn_kernel_prune_ref (kernel_ref_prune.cpp:39-123):
 * the list as searched (at rlist) pruned to the j-clusters with at least one
 * atom pair closer than rlist_inner with the current coordinates (those of
 * the last put_on_grid, convert_coordinates or compute), keeping their order
 * and exclusion masks and dropping emptied i-entries.  Every prune starts
 * from the list as searched; the next construct_pairlist replaces it.
 * rcoulomb <= rlist_inner <= rlist. */
int nbx_prune_pairlist(nbx_context* context, double rlist_inner);

/* Dynamic pruning inside the step loop, rolling over parts
 * (nonbonded_verlet_t::dispatchPruneKernelCpu/Gpu, nbnxm.h:306-316, with
 * the list split into numRollingPruningParts parts, pairlistparams.h:136-138,
 * and the outer list kept as pairlist.cpp:4264-4283 keeps it; the prune
 * criterion of kernel_ref_prune.cpp:39-123): the pair kernel's work layout,
 * as built from the list as searched, is pruned to the j-clusters with at
 * least one atom pair closer than rlist_inner with the current coordinates
 * (those of the last put_on_grid, convert_coordinates or compute), for the
 * kernel wavefronts w with w % num_parts == part; the other parts keep their
 * last prune.  Every prune starts from the layout as built.  The layout's
 * groups, window slots, schedule and staging stay, so nothing is rebuilt,
 * nothing is allocated after the first call and the host does not wait:
 * the call only queues three kernels.  nbx_get_pairlist still returns the
 * list as searched; nbx_get_work_groups shows the pruned layout.
 * rcoulomb <= rlist_inner <= rlist, 0 <= part < num_parts. */
int nbx_prune_rolling(nbx_context* context, double rlist_inner, int part, int num_parts);

/* nonbonded_verlet_t::convertCoordinates (nbnxm.cpp:129-141). */
int nbx_convert_coordinates(nbx_context* context, const float* x, int x_on_device);

/* nonbonded_verlet_t::dispatchNonbondedKernel (kerneldispatch.cpp:420-456)
 * with enbvClearFYes: clears the cluster force buffer, then runs the
 * cluster-pair kernel; flags select energies and shift forces. */
int nbx_dispatch_kernel(nbx_context* context, unsigned flags);

/* nonbonded_verlet_t::atomdata_add_nbat_f_to_f (nbnxm.cpp:162-180) after
 * nblib's zero fill (gmxcalculator.cpp:76-82): f is overwritten with the
 * nonbonded forces in the caller's atom order. */
int nbx_add_nbat_f_to_f(nbx_context* context, float* f, int f_on_device);

/* GmxForceCalculator::compute (gmxcalculator.cpp:70-83): convert
 * coordinates, dispatch the kernel, reduce forces into f (overwritten, or
 * with NBX_COMPUTE_ADD_FORCES added to what f holds).
 * With NBX_COMPUTE_VIRIAL the virial is also formed (calc_vir,
 * calcvir.cpp:103-148).  Device pointers are used asynchronously on the
 * context's stream; host pointers make the call synchronous. */
int nbx_compute(nbx_context* context, const float* x, float* f, unsigned flags, int on_device);

/* nonbonded_verlet_t::setAtomProperties (nbnxm.h:268-272; charges into the
 * atom data, atomdata.cpp:784-836) for new per-atom charges without
 * re-gridding: the grid and pair list stay valid. */
int nbx_set_charges(nbx_context* context, const float* charges, int on_device);

/* Adjoint of the forces (no counterpart in the reference, whose kernels
 * are forward only): for a loss L(F) with g = dL/dF (num_atoms x 3, atom
 * order), at coordinates x and on the current pair list (recomputed pair by
 * pair, nothing stored from the forward call):
 *   grad_x   dL/dx, num_atoms x 3 floats (the Hessian-vector product H g);
 *   grad_q   dL/dq, num_atoms doubles;
 *   grad_c6, grad_c12  dL/dC6 and dL/dC12 for the matrices given to
 *            nbx_set_nonbonded_parameters, num_types^2 doubles each, every
 *            entry taken as independent (for a symmetric parameter
 *            C_ij = C_ji, i != j, the derivative is grad[i][j] + grad[j][i]).
 * Any output may be NULL (no C6/C12 output skips their accumulation).  All
 * pointers are device pointers, used asynchronously on the context's
 * stream, when on_device is set; host pointers make the call synchronous.
 * At most 31 particle types; the LJ cut-off with the potential shift only
 * (a context made with another vdw_modifier or LJ-PME gets
 * NBX_ERROR_INPUT). */
int nbx_compute_adjoint(nbx_context* context, const float* x, const float* g, float* grad_x, double* grad_q,
                        double* grad_c6, double* grad_c12, int on_device);

/* Energies of the last compute with NBX_COMPUTE_ENERGY (gmx_enerdata_t
 * terms reduced by reduce_energies_over_lists, kernel_common.cpp:109-133):
 * LJ and Coulomb (real-space) in kJ/mol.  Synchronises. */
int nbx_get_energies(nbx_context* context, double* v_lj, double* v_coulomb);

/* Shift forces of the last compute (nbnxn_atomdata_add_nbat_fshift_to_fshift,
 * atomdata.cpp:1471-1487): 45 x 3 doubles in XYZ2IS order (ishift.h:40-55). */
int nbx_get_shift_forces(nbx_context* context, double* fshift);

/* Virial of the last compute with NBX_COMPUTE_VIRIAL: 3 x 3 doubles,
 * -1/2 sum x_i (x) f_i - 1/2 sum s (x) fshift_s (sim_util.cpp:157-177). */
int nbx_get_virial(nbx_context* context, double* virial);

/* Waits for the context's queued work. */
int nbx_synchronize(nbx_context* context);

/* Grid and list sizes (the reference prints these via print_nblist_statistics,
 * pairlist.cpp:2955-3018). */
int nbx_get_pairlist_info(nbx_context* context, nbx_pairlist_info* info);

/* The grid's atom order: atom_indices[4 * num_clusters] (-1 = filler) and
 * cells[num_atoms] (GridSetData atomIndices / cells, gridsetdata.h). */
int nbx_get_grid(nbx_context* context, int* atom_indices, int* cells);

/* The cluster-pair list: i_entries[4 * num_i_entries] {ci, shift, cj_start,
 * cj_end} and j_entries[2 * num_j_entries] {cj, exclusion mask}
 * (NbnxnPairlistCpu ci/cj arrays, pairlist.h:226-259). */
int nbx_get_pairlist(nbx_context* context, int* i_entries, uint32_t* j_entries);

/* The atom pairs the list holds with squared distance below r^2, computed
 * on the device with the canonical float rounding, as keys
 * (min atom << 32 | max atom), unsorted.  Returns the number of pairs (or a
 * negative error); writes at most cap. */
int64_t nbx_list_pairs(nbx_context* context, double r, int64_t* out, int64_t cap);

/* No counterpart in the reference: nbx_list_pairs for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid hipMalloc).  The same pair set, written in list order (i-entry,
 * j-entry, i-atom, j-atom) to `out`, a DEVICE array of cap keys; returns the
 * number of pairs (writes at most cap) or a negative error.  Nothing is
 * allocated: per-i-entry counts and offsets live in a static device array of
 * 2 MiB that each GPU it runs on keeps until the process ends, so lists of
 * at most 262,144 i-entries are taken (larger ones are an input error).
 * Calls on one GPU share that array: calls that overlap on one GPU race.
 * The host waits for the result (one 4-byte read). */
int64_t nbx_list_pairs_noalloc(nbx_context* context, double r, int64_t* out, int64_t cap);

/* The kernel's work groups (no counterpart in the reference, whose CPU
 * kernels need none): info[5] = {group size G, number of groups, number of
 * chunks, LDS window slots, work-list length}; groups[4 * num_groups] =
 * {first chunk, end chunk, union offset, union size}; chunks[4 * num_chunks]
 * = {ci, shift | self << 8 | leading masked tiles << 9, work start, tiles of
 * 16 entries}; union_list[sum of union sizes] (j-clusters of each group other
 * than its own G); work[2 * work-list length] = {cj, window slot << 16 |
 * mask}, padding entries {num_clusters, own slot << 16 | 0xffff}.  Pointers
 * may be NULL except info. */
int nbx_get_work_groups(nbx_context* context, int* info, int* groups, int* chunks, int* union_list,
                        uint32_t* work);

/* Records HIP events on the context's stream around each of the next
 * max_launches launches of the cluster-pair kernel (0 stops recording); no
 * counterpart in the reference, which times its CPU kernels with cycle
 * counters (nbnxm/benchmark/bench_setup.cpp:300-330).  For benchmarking. */
int nbx_set_kernel_timing(nbx_context* context, int max_launches);

/* Average duration in ms of the recorded pair-kernel launches and their
 * number; waits for them, then starts a new record. */
int nbx_get_kernel_timing(nbx_context* context, double* average_ms, int* launches);

/* beta, table scale, table size, epsfac, k_rf, c_rf, sh_ewald, LJ
 * dispersion and repulsion potential shifts (interaction_const_t). */
int nbx_get_interaction_constants(nbx_context* context, double* out9);

/* The LJ part of interaction_const_t as the kernels use it: rvdw, the switch
 * radius, the force switch's dispersion and repulsion c2, c3
 * (force_switch_constants, forcerec.cpp:787-801), the potential switch's c3,
 * c4, c5 (potential_switch_constants :803-818), and LJ-PME's beta_lj^2,
 * beta_lj^6 / 6 and sh_lj_ewald (forcerec.cpp:703-717): 12 values. */
int nbx_get_vdw_constants(nbx_context* context, double* out12);

/* The Ewald correction table as the kernel reads it (interaction_const_t's
 * coulombEwaldTables->tableFDV0, generateEwaldCorrectionTables,
 * tables/forcetable.cpp:144-298): *num_points quadruplets {F_i, F_i+1 - F_i,
 * V_i, V_i + sh_ewald} in float (the reference leaves the fourth entry 0).
 * With fdv0 NULL only *num_points is set (0 without Ewald). */
int nbx_get_ewald_table(nbx_context* context, int* num_points, float* fdv0);

/* ------------------------------------------------------------------------
 * Batched ensemble: many independent systems (replicas of uneven size) that
 * share the options and LJ parameters, computed by one launch of the pair
 * kernel over all their work groups.  The reference runs an ensemble as one
 * simulation per member (gmxapi's ensemble operations,
 * python_packaging/src/gmxapi/operation.py, simulation/mdrun.py:162-200);
 * nblib computes one system per ForceCalculator.  Coordinates and forces of
 * the whole batch are device arrays with the systems' atoms concatenated in
 * the order they were added.  Shift forces and the virial are per-system
 * calls (nbx_compute on a single context); the batch computes forces and
 * energies.
 * ------------------------------------------------------------------------ */
typedef struct nbx_batch nbx_batch;

/* c6 and c12 as for nbx_set_nonbonded_parameters. */
int nbx_batch_create(const nbx_options* options, int num_types, const double* c6, const double* c12,
                     nbx_batch** batch);
int nbx_batch_destroy(nbx_batch* batch);
int nbx_batch_set_stream(nbx_batch* batch, void* stream);

/* Adds one system (host arrays, as nbx_set_topology and nbx_put_on_grid
 * take them) and returns its index. */
int nbx_batch_add_system(nbx_batch* batch, int num_atoms, const int* types, const float* charges,
                         const int* exclusion_offsets, const int* exclusion_indices, const float* box,
                         const float* x, int* index);

/* Grids and searches every system (as nbx_put_on_grid + nbx_construct_pairlist)
 * and concatenates their kernel work into one launch. */
int nbx_batch_construct_pairlists(nbx_batch* batch);

/* {systems, atoms, clusters, cluster pairs (j-entries), kernel work groups,
 * group size, workgroups per CU} */
int nbx_batch_get_info(nbx_batch* batch, int64_t* info7);

/* Forces (and with NBX_COMPUTE_ENERGY energies) of every system: x and f are
 * device arrays of 3 * total atoms floats; f is overwritten.  Queued on the
 * batch's stream. */
int nbx_batch_compute(nbx_batch* batch, const float* x, float* f, unsigned flags);

/* Records one nbx_batch_compute with these arguments into a HIP graph
 * (replaces any earlier one), and replays it. */
int nbx_batch_capture_graph(nbx_batch* batch, const float* x, float* f, unsigned flags);
int nbx_batch_launch_graph(nbx_batch* batch);

/* {LJ, Coulomb} per system (2 * systems doubles) of the last compute with
 * energies.  Synchronises. */
int nbx_batch_get_energies(nbx_batch* batch, double* energies);

/* Times the candidate kernel schedules (work-group size, workgroups per CU,
 * cost weight of an i-entry in the static schedule, in pair-list entries) on
 * this batch with `repeats` launches each and keeps the fastest;
 * chosen[3] = {group size, workgroups per CU, i-entry weight}. */
int nbx_batch_autotune(nbx_batch* batch, const float* x, float* f, unsigned flags, int repeats, int* chosen3);

/* The per-system contexts (for sequential reference launches and checks);
 * owned by the batch. */
int nbx_batch_get_system(nbx_batch* batch, int index, nbx_context** context);

#ifdef __cplusplus
}
#endif

#endif /* NBX_H */
