/*
 * sponge_nb.h -- C ABI of the gfx950 nonbonded-force library (libsponge_nb.so).
 *
 * The library replaces these modules of SPONGE 2.0 on one GPU:
 *   NEIGHBOR_LIST              SPONGE/neighbor_list/neighbor_list.{h,cpp}
 *   LENNARD_JONES_INFORMATION  SPONGE/Lennard_Jones_force/Lennard_Jones_force.{h,cpp}
 *   Particle_Mesh (PME)        SPONGE/PM_force/PM_force.{h,cpp}
 *   NON_BOND_14                SPONGE/nb14/nb14.{h,cpp}
 *   BOND                       SPONGE/bond/bond.{h,cpp}
 *   NVE leap-frog, MIDDLE_Langevin_INFORMATION, SETTLE
 *                              SPONGE/MD_core/nve.hpp, SPONGE/thermostat/Middle_Langevin_MD.cpp,
 *                              SPONGE/constrain/settle.cpp
 *   stress / pressure assembly SPONGE/MD_core/sys.hpp, SPONGE/MD_core/MD_core.cpp
 * Each entry point below names the reference method it stands in for; the
 * argument lists keep the reference's order and meaning, with the module
 * object replaced by an snb_context.
 *
 * Conventions
 *  - Pointers named d_* are device pointers owned by the caller (as the
 *    reference's dd.* buffers are, Domain_decomposition.h:24-72); h_* are
 *    host pointers read during the call.  Every launch goes to the context's
 *    stream; calls return without synchronising unless stated.
 *  - Forces, energies and virials are accumulated (+=) into caller buffers,
 *    as in the reference.  Internally every term is summed in int64 fixed
 *    point (scale 2^32) and converted once at the boundary.  Passing NULL for
 *    an output keeps the contribution in the context's fixed-point
 *    accumulators until snb_flush_forces / snb_flush_atom_energy.
 *  - Every function returns an SNB_* status and never exits the process
 *    (the reference calls Throw_SPONGE_Error -> exit, utils/control/error.hpp:30).
 *  - One rank only: local order = global order, ghost_numbers must be 0,
 *    solvent_numbers must be 0 (this library owns the water atoms too, in
 *    place of SOLVENT_LENNARD_JONES, Lennard_Jones_force/solvent_LJ.cpp).
 */
#ifndef SPONGE_NB_H
#define SPONGE_NB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Value types with the reference's memory layout. */
typedef struct { float x, y, z; } snb_vector;                          /* VECTOR, utils/vector.hpp:30-33 */
typedef struct { float a11, a21, a22, a31, a32, a33; } snb_ltmatrix3;  /* LTMatrix3, utils/vector.hpp:256-258 */
typedef struct { int atom_numbers; int ghost_numbers; int* atom_serial; } snb_atom_group; /* ATOM_GROUP, common.h:175-180 */

typedef struct snb_context snb_context;

/* return codes (reference error codes: utils/control/error.hpp:4-27) */
#define SNB_OK 0
#define SNB_ERR_ARGUMENT 1      /* bad argument value or NULL where required */
#define SNB_ERR_HIP 2           /* a HIP runtime call failed */
#define SNB_ERR_BOX_TOO_SMALL 3 /* fewer than 3 NL cells along an axis (neighbor_list.cpp:809-814) */
#define SNB_ERR_UNSUPPORTED 4   /* ghosts, solvent split, FFT size with a prime > 7, ... */
#define SNB_ERR_NOT_INITIALIZED 5
#define SNB_ERR_OUT_OF_MEMORY 6
#define SNB_ERR_DEVICE_STATUS 7 /* a sticky device flag is set, see snb_read_status */
#define SNB_ERR_TOO_LARGE 8     /* atom_numbers above SNB_NOALLOC_MAX_ATOMS (allocation-free entry point) */

/* sticky device flags (snb_read_status); the reference keeps the first three
 * as device ints but never reads them (neighbor_list.cpp:887-965) */
#define SNB_FLAG_NL_OVERFLOW 1      /* a row exceeded max_neighbor_numbers; row truncated */
#define SNB_FLAG_CELL_OVERFLOW 2    /* a cell held more than max_atom_in_grid_numbers atoms */
#define SNB_FLAG_NONFINITE_CRD 4    /* a coordinate was NaN/Inf (PME uses 0 for it, PM_force.cpp:710-714) */

int snb_abi_version(void);
const char* snb_status_string(int code);

/* Creates a context for atom_numbers atoms on HIP device `device`, issuing
 * work on hip_stream (hipStream_t; NULL = the null stream). */
int snb_context_create(snb_context** out, int device, void* hip_stream, int atom_numbers);
int snb_context_destroy(snb_context* ctx);
int snb_set_stream(snb_context* ctx, void* hip_stream);
/* Runs the PME reciprocal pipeline on a second, library-owned stream beside
 * the direct-space work (default 1).  It starts once the step's coordinates
 * are final -- at the last snb_nl_update, or at the reciprocal call when no
 * update came since the previous one -- and every later call that reads its
 * results (force flush, copies, energies, status) waits for it on the
 * context stream.  With overlap on, crd and charge must not change between a
 * step's snb_nl_update and its snb_pme_reciprocal_force, as in the
 * reference's step loop.  Results are identical either way (no counterpart
 * in the reference, which runs everything on the null stream). */
int snb_set_pme_overlap(snb_context* ctx, int enable);
int snb_synchronize(snb_context* ctx);
/* Reads (and with clear != 0, resets) the sticky SNB_FLAG_* bits; synchronises. */
int snb_read_status(snb_context* ctx, int clear, int* flags);

/* ---------------------------------------------------------------- exclusions */
/* Uploads the excluded-atom CSR of MD_INFORMATION::non_bond_information
 * (MD_core/nb.hpp:30-70): for atom i, h_excluded_list[h_excluded_list_start[i]
 * .. + h_excluded_numbers[i]) lists its excluded partners (the native
 * exclude file lists j > i).  The same arrays are what the reference passes as
 * device pointers to NEIGHBOR_LIST::Update and the PME exclusion correction;
 * those entry points below also accept them directly. */
int snb_set_exclusions(snb_context* ctx, const int* h_excluded_list_start,
                       const int* h_excluded_list, const int* h_excluded_numbers,
                       int excluded_atom_numbers);

/* ------------------------------------------------------------ neighbor list */
/* NEIGHBOR_LIST::Initial (neighbor_list.cpp:729-832) with the controller keys
 * resolved by the caller: neighbor_list_max_neighbor_numbers (default 1200),
 * neighbor_list_max_atom_in_grid_numbers (150), neighbor_list_refresh_interval
 * (0 = displacement criterion), neighbor_list_skin_permit (0.5).  Cells are
 * (cutoff+skin)/2 wide; boxes with < 3 cells on an axis return
 * SNB_ERR_BOX_TOO_SMALL.  The windowed pair path (below) is chosen only when
 * max_atom_in_grid_numbers <= 256; a larger capacity keeps the plain path,
 * whose cells hold up to max_atom_in_grid_numbers atoms as the reference's.
 * A triclinic cell takes the plain path with the reference's neighbour-grid
 * table (Find_Neighor_Grids_Device, :5-48, tilts truncated to whole grids):
 * the list holds the pairs that table pairs up, which can leave out pairs
 * within cutoff + skin for tilts below a grid width, as the reference's does;
 * SNB_ERR_UNSUPPORTED when a grid would have more than 192 neighbour grids. */
int snb_nl_initial(snb_context* ctx, float cutoff, float skin, snb_ltmatrix3 cell,
                   snb_ltmatrix3 rcell, int max_neighbor_numbers,
                   int max_atom_in_grid_numbers, int refresh_interval,
                   float skin_permit);

/* NEIGHBOR_LIST::Update (neighbor_list.cpp:834-885).  update = 1 forces a
 * rebuild (FORCED_UPDATE); otherwise the reference policy applies: with a
 * positive refresh interval the step number decides on the host and a step
 * without a rebuild launches nothing; with interval 0 the displacement test
 * runs on the device and gates the rebuild there (no host synchronisation).
 * atom_local may be
 * NULL (identity); ghost_numbers must be 0.  The exclusion arrays may be NULL
 * when snb_set_exclusions was called.  The cell may differ from the initial
 * one (barostat steps): cells keep the initial width and the last cell of an
 * axis takes the remainder, as in the reference; a box smaller than the
 * initial cell count times that width returns SNB_ERR_BOX_TOO_SMALL
 * (re-initialise, as Main_Box_Change_Largely does). */
int snb_nl_update(snb_context* ctx, const int* d_atom_local, int local_atom_numbers,
                  int ghost_numbers, const snb_vector* d_crd, snb_ltmatrix3 cell,
                  snb_ltmatrix3 rcell, int step, int update,
                  const int* d_excluded_list_start, const int* d_excluded_list,
                  const int* d_excluded_numbers);

/* The public half list (NEIGHBOR_LIST::d_nl, neighbor_list.h:42-45): an
 * ATOM_GROUP per atom whose atom_serial points into a padded
 * int[atom_numbers][max_neighbor_numbers] device table; row i holds j > i. */
int snb_nl_device_list(snb_context* ctx, snb_atom_group** d_nl, int** d_serial_table,
                       int* max_neighbor_numbers);
/* Copies the row lengths (int[atom_numbers]) and the padded table
 * (int[atom_numbers][max_neighbor_numbers]) of the public half list into
 * caller device buffers (the reference's h_nl readback, neighbor_list.h:42). */
int snb_nl_copy_rows(snb_context* ctx, int* d_counts, int* d_rows);
/* Whether snb_nl_update also writes the public list above (no counterpart
 * in the reference): 1 (default) = yes, the reference's behaviour, with the
 * entries of a row in an order that can differ between rebuilds (as the
 * reference's, whose cell binning appends with atomics); 2 = yes, with every
 * row sorted ascending (deterministic, about 1 ms more per rebuild at 93k
 * atoms); 0 = no.  The library's LJ kernel does not read it: with every axis
 * >= 5 cells it consumes an internal cell-ordered list of 16-bit window slots
 * with the same pair set, and its integer force sums do not depend on the
 * order of a row.  Callers that never read d_nl may switch the export off. */
int snb_nl_set_export(snb_context* ctx, int enable);
/* The internal list in atom ids: d_counts[i] and the row
 * d_rows[i * max_neighbor_numbers ..] of every pair whose internal row owner
 * is atom i (each pair once; owner = the atom whose cell precedes by the
 * forward cell rule).  Equals the public list when the windowed path is off. */
int snb_nl_copy_internal(snb_context* ctx, int* d_counts, int* d_rows);
/* Per-atom energy and virial of a pair (atom_energy, atom_direct_cf_energy,
 * atom_virial of snb_lj_pme_direct_force): by default they go to the lower atom
 * id, the reference's row atom, so the per-atom arrays equal the reference's.
 * enable = 1 keeps them with the windowed list's row owner instead: the same
 * totals with fewer atomics on energy/virial steps (no counterpart in the
 * reference). */
int snb_lj_set_owner_attribution(snb_context* ctx, int enable);
/* LDS slots the windowed pair kernel reserves per block (default 1.3x the
 * mean window, at most what the device's LDS holds).  Blocks whose window is
 * larger take a slower path with global integer atomics; results are the same.
 * A smaller reservation lets more work of the reciprocal stream share a
 * compute unit (no counterpart in the reference). */
int snb_nl_set_window_capacity(snb_context* ctx, int slots);
/* Block shape of the windowed pair path: each workgroup owns a x b x s cells
 * (x, y, z); 0, 0, 0 selects the plain path (one row per thread group over
 * the public list, global integer atomics).  The default comes from a fixed
 * preference list; a tuner can time the shapes this call accepts and keep the
 * fastest.  SNB_ERR_ARGUMENT when the grid cannot take the shape or its
 * window does not fit the LDS.  Synchronises; the next snb_nl_update
 * rebuilds (setting the shape in use changes nothing).  Forces, energies and
 * virials are the same for every shape (no counterpart in the reference). */
int snb_nl_set_window_shape(snb_context* ctx, int a, int b, int s);
/* How the windowed pair kernel hands its window forces to the accumulators
 * on force-only steps: 0 = through a per-block stage and a per-atom gather
 * kernel, 1 = atomic integer adds at the end of each block, -1 (default) =
 * by size (1 below 65,536 atoms and once the stage would exceed 128 MiB,
 * about 310k atoms; 0 between).  Energy
 * and virial steps always use the stage.  Results are the same either way
 * (integer sums); SNB_ERR_ARGUMENT for another mode (no counterpart in the
 * reference). */
int snb_lj_set_direct_j(snb_context* ctx, int mode);
/* The windowed pair kernel's loop on force-only steps: 0 = a software
 * pipeline within each list row, 1 = one pipeline per wavefront across the
 * ends of its rows (faster alone; it holds more registers, which leaves
 * fewer slots for the reciprocal pipeline's workgroups beside it), -1
 * (default) = by size (1 from 262,144 atoms).  Results are the same bits
 * either way; SNB_ERR_ARGUMENT for another mode (no counterpart in the
 * reference). */
int snb_lj_set_row_stream(snb_context* ctx, int mode);
/* Force-only steps of the row kernel (the plain path, used when the windowed
 * path is off) for systems of at most 2,560 atoms: 1 = every atom's force
 * accumulators in LDS, one workgroup per 64 rows and one global add per
 * nonzero sum; 0 = global integer atomics per pair; -1 (default) = LDS only
 * while recording a step batch, where several small systems share the
 * device and the global atomics' throughput bounds the kernel (alone it is
 * bound by their latency, and the two measure the same).  Results are the
 * same either way (no counterpart in the reference). */
int snb_lj_set_row_lds(snb_context* ctx, int mode);
/* Geometry of the windowed pair path (windowed = 0 when not used). */
int snb_nl_window_info(snb_context* ctx, int* windowed, int* a, int* b, int* s, int* wcap,
                       int* nblock);
/* Cell grid shape chosen by snb_nl_initial (NEIGHBOR_LIST::GRIDS Nx/Ny/Nz). */
int snb_nl_grid_shape(snb_context* ctx, int* nx, int* ny, int* nz);

/* --------------------------------------------------------------------- LJ */
/* LENNARD_JONES_INFORMATION::Initial (Lennard_Jones_force.cpp:186-283) given
 * the loaded tables: h_LJ_A/h_LJ_B are the in-memory pair tables
 * (file value x12 and x6, xponge/load/native/lj.hpp:47,57), length
 * atom_type_numbers*(atom_type_numbers+1)/2; h_atom_LJ_type per atom.  Also
 * computes the long-range correction factor (:246-272) from a type histogram
 * instead of the O(N^2) Total_C6_Get. */
int snb_lj_initial(snb_context* ctx, int atom_type_numbers, const float* h_LJ_A,
                   const float* h_LJ_B, const int* h_atom_LJ_type, float cutoff);

/* LENNARD_JONES_INFORMATION::LJ_PME_Direct_Force_With_Atom_Energy_And_Virial
 * (Lennard_Jones_force.cpp:359-417; declaration Lennard_Jones_force.h:138-144).
 * LJ + erfc-screened Coulomb over the half list d_nl for pairs with r < cutoff.
 * d_frc += forces; with need_atom_energy: d_atom_energy += LJ + Coulomb per
 * atom and d_atom_direct_pme_energy = Coulomb per atom; with need_virial:
 * d_atom_virial += per-atom virial (reference convention: -f (x) dr, with the
 * off-diagonals symmetrised, vector.hpp:336-347).  Any output may be NULL
 * (then kept in the fixed-point accumulators).  SNB_ERR_ARGUMENT for
 * pme_beta < 0 or pme_beta * cutoff > 4.9 (an Ewald tolerance below 1e-11);
 * SNB_ERR_UNSUPPORTED when pme_beta changes while the stream is being
 * captured into a graph (the Coulomb force table is rebuilt on a change). */
int snb_lj_pme_direct_force(snb_context* ctx, int atom_numbers, int local_atom_numbers,
                            int solvent_numbers, int ghost_numbers,
                            const snb_vector* d_crd, const float* d_charge,
                            snb_vector* d_frc, snb_ltmatrix3 cell, snb_ltmatrix3 rcell,
                            const snb_atom_group* d_nl, float pme_beta,
                            int need_atom_energy, float* d_atom_energy,
                            int need_virial, snb_ltmatrix3* d_atom_virial,
                            float* d_atom_direct_pme_energy);

/* LENNARD_JONES_INFORMATION::Long_Range_Correction (Lennard_Jones_force.cpp:320-341):
 * d_virial[0] diagonal += 2 f/V, d_potential[0] += f/V. */
int snb_lj_long_range_correction(snb_context* ctx, int need_pressure, snb_ltmatrix3* d_virial,
                                 int need_potential, float* d_potential, float volume);

/* LENNARD_JONES_INFORMATION::Step_Print (Lennard_Jones_force.cpp:419-433):
 * LJ_short = sum of per-atom LJ energies of the last energy call (exact
 * int64 sum), LJ_long = f/V.  Synchronises. */
int snb_lj_energy(snb_context* ctx, double* lj_short, double* lj_long);

/* -------------------------------------------------------------------- 1-4 */
/* NON_BOND_14::Initial (nb14.cpp:103-185) with the pair list resolved by the
 * caller the way Xponge::Native_Load_NB14 does (xponge/load/native/nb14.hpp:8-95):
 * per pair the atoms a and b, A in the x12 form and B in the x6 form (the LJ
 * scale already applied), and the Coulomb scale.  nb14_numbers = 0 leaves the
 * module uninitialised and its force call a no-op, as in the reference. */
int snb_nb14_initial(snb_context* ctx, int nb14_numbers, const int* h_atom_a,
                     const int* h_atom_b, const float* h_A, const float* h_B,
                     const float* h_cf_scale_factor);

/* NON_BOND_14::Non_Bond_14_LJ_CF_Force_With_Atom_Energy_And_Virial
 * (nb14.cpp:284-301, kernel :7-100): for each pair, with dr = pbc(r_b - r_a),
 * f = -A r^-14 + B r^-8 - cf q_a q_b r^-3 (charges below 1e-10 in magnitude
 * raised to 1e-10); d_frc[a] += f dr, d_frc[b] -= f dr; with
 * need_atom_energy: d_atom_energy[a] += cf q_a q_b / r + A/12 r^-12 - B/6 r^-6
 * (the module's per-atom energy restarts each energy call); with need_virial:
 * d_atom_virial[a] += -f dr (x) dr.  NULL outputs stay in the fixed-point
 * accumulators. */
int snb_nb14_force(snb_context* ctx, const snb_vector* d_crd, const float* d_charge,
                   snb_ltmatrix3 cell, snb_ltmatrix3 rcell, snb_vector* d_frc,
                   int need_atom_energy, float* d_atom_energy, int need_virial,
                   snb_ltmatrix3* d_atom_virial);

/* NON_BOND_14::Step_Print (nb14.cpp:303-324): the sums of the 1-4 LJ
 * (nb14_LJ) and Coulomb (nb14_EE) pair energies of the last energy call, as
 * exact int64 sums.  Synchronises. */
int snb_nb14_energy(snb_context* ctx, double* lj14, double* ee14);

/* ------------------------------------------------------------------- bonds */
/* BOND::Initial (bond/bond.cpp:66-157) with the list read by the caller
 * (xponge/load/native/bond.hpp: "N" then "a b k r0" per bond; E = k (r - r0)^2).
 * bond_numbers = 0 leaves the module uninitialised (calls are no-ops). */
int snb_bond_initial(snb_context* ctx, int bond_numbers, const int* h_atom_a,
                     const int* h_atom_b, const float* h_k, const float* h_r0);

/* BOND::Bond_Force_With_Atom_Energy_And_Virial (bond.cpp:243-260, kernel
 * :15-64): with dr = pbc(r_a - r_b), f = -2 k (r - r0) dr / r on a and -f on b;
 * d_atom_energy[a] += k (r - r0)^2 (the module's per-atom energy restarts each
 * energy call); d_atom_virial[a] += f (x) dr.  NULL outputs stay in the
 * fixed-point accumulators. */
int snb_bond_force(snb_context* ctx, const snb_vector* d_crd, snb_ltmatrix3 cell,
                   snb_ltmatrix3 rcell, snb_vector* d_frc, int need_atom_energy,
                   float* d_atom_energy, int need_virial, snb_ltmatrix3* d_atom_virial);

/* BOND::Step_Print (bond.cpp:262-277): the sum of the bond energies of the
 * last energy call (exact int64 sum).  Synchronises. */
int snb_bond_energy(snb_context* ctx, double* energy);

/* ------------------------------------------------------------- integration */
/* MD_INFORMATION::NVE_iteration::Leap_Frog (MD_core/nve.hpp:3-48):
 * v += dt inv_m f, x += dt v, with |v| capped at max_velocity when it is > 0
 * (velocity_max).  dt in the internal time unit (ps * 20.455). */
int snb_leap_frog(snb_context* ctx, const float* d_inverse_mass, const snb_vector* d_frc,
                  snb_vector* d_vel, snb_vector* d_crd, float dt, float max_velocity);

/* MIDDLE_Langevin_INFORMATION::Initial (thermostat/Middle_Langevin_MD.cpp:137-245)
 * with the controller keys resolved by the caller: target_temperature (K),
 * thermostat_tau (ps; friction 1/tau), thermostat_seed, dt (ps),
 * velocity_max (<= 0: none).  Masses in amu; a zero mass gets neither force
 * nor noise.  The noise comes from the library's counter-based generator
 * (Philox4x32-10 + Box-Muller) seeded with `seed`, not from curand: the
 * sequence differs from the reference's, its distribution does not. */
int snb_langevin_initial(snb_context* ctx, const float* h_mass, float target_temperature,
                         float tau, int seed, float dt, float max_velocity);

/* MIDDLE_Langevin_INFORMATION::MD_Iteration_Leap_Frog
 * (Middle_Langevin_MD.cpp:305-331, kernel :5-66): per atom
 * v1 = v + dt f/m; x1 = x + dt/2 v1; v = e^{-gamma dt} v1 + sqrt(kT(1 - e^{-2 gamma dt})/m) R;
 * x = x1 + dt/2 v.  d_acc is accepted for the reference's signature and unused. */
int snb_langevin_iterate(snb_context* ctx, const snb_vector* d_frc, snb_vector* d_vel,
                         snb_vector* d_acc, snb_vector* d_crd);

/* The thermostat's dt (internal unit) and e^{-gamma dt}, and the constraint
 * factors main.cpp:650-652 sets after the iteration: v_factor = e^{-gamma dt},
 * x_factor = (e^{-gamma dt} + 1)/2.  NULL outputs are skipped. */
int snb_langevin_parameters(snb_context* ctx, float* dt_internal, float* exp_gamma,
                            float* v_factor, float* x_factor);

/* SETTLE::Initial (constrain/settle.cpp:342-571) for rigid triangles: the
 * caller lists each triangle (A = the atom bonded to both others, B, C) with
 * its three constraint lengths, as the reference finds them among the
 * constraint pairs.  Two-atom rigid pairs: snb_settle_pair_initial.  Zero
 * triangles and zero pairs leave SETTLE uninitialised (its calls are no-ops). */
int snb_settle_initial(snb_context* ctx, int triangle_numbers, const int* h_atom_A,
                       const int* h_atom_B, const int* h_atom_C, const float* h_rAB,
                       const float* h_rAC, const float* h_rBC, const float* h_mass);
/* The rigid two-atom pairs of SETTLE::Initial (settle.cpp:342-571, CONSTRAIN_PAIR):
 * atoms i and j at distance constant_r, constrain_k = 1/(m_i + m_j) from
 * h_mass, as the reference finds them among the constraint pairs (no atom
 * in two pairs, SNB_ERR_ARGUMENT otherwise).  snb_settle_remember keeps each
 * pair's j - i and snb_settle_do applies settle_pair (:258-307) before the
 * triangles, its virial included when pressure is asked for.  Zero pairs
 * removes them.  The fused iteration (snb_md_iterate_fused) and the FP64 step
 * take triangles only: SNB_ERR_UNSUPPORTED while pairs are set. */
int snb_settle_pair_initial(snb_context* ctx, int pair_numbers, const int* h_atom_i,
                            const int* h_atom_j, const float* h_constant_r, const float* h_mass);

/* SETTLE::Remember_Last_Coordinates (settle.cpp:573-640): B - A and C - A of
 * every triangle before the iteration moves the atoms. */
int snb_settle_remember(snb_context* ctx, const snb_vector* d_crd, snb_ltmatrix3 cell,
                        snb_ltmatrix3 rcell);

/* SETTLE::Do_SETTLE (settle.cpp:659-697, kernels :101-253): restores each
 * triangle's geometry and corrects the velocities by
 * v_factor / dt / x_factor times the position change; with need_pressure,
 * d_stress[0] += (1/V) * the constraint virial.  dt (internal unit),
 * v_factor and x_factor are the CONSTRAIN module's (dt, v_factor, x_factor). */
int snb_settle_do(snb_context* ctx, const float* d_mass, snb_vector* d_crd, snb_ltmatrix3 cell,
                  snb_ltmatrix3 rcell, snb_vector* d_vel, int need_pressure,
                  snb_ltmatrix3* d_stress, float dt, float v_factor, float x_factor);

/* Fused iteration (no counterpart in the reference, whose main.cpp:634-712
 * makes these separate calls): SETTLE's Remember_Last_Coordinates, the middle
 * Langevin update (MIDDLE_Langevin_INFORMATION::MD_Iteration_Leap_Frog) and
 * Do_SETTLE with the thermostat's constraint factors, one thread per rigid
 * triangle or free atom, with the forces taken from the context's fixed-point
 * accumulators (the force calls of the step ran with d_frc = NULL) and
 * cleared.  The result is bit-identical to snb_flush_forces into a zeroed
 * force array followed by the three separate calls; d_frc is not written.
 * Needs snb_langevin_initial; SNB_ERR_UNSUPPORTED when triangles share atoms.
 * For callers that drive whole steps and do not read the float forces. */
int snb_md_iterate_fused(snb_context* ctx, const float* d_mass, snb_vector* d_crd,
                         snb_ltmatrix3 cell, snb_ltmatrix3 rcell, snb_vector* d_vel,
                         int need_pressure, snb_ltmatrix3* d_stress);

/* --------------------------------------------------------------- pressure */
/* The stress of a pressure step is assembled as the reference does it: the
 * caller zeroes d_virial_tensor and d_stress at the start of the step
 * (MD_Reset_Atom_Energy_And_Virial_And_Force, MD_core.cpp:654-676), SETTLE adds its constraint
 * virial, then these calls add the potential and kinetic parts.  Atom sums are
 * exact (each float entry, or each FP64 product m v_a v_b, rounded to int64
 * at 2^32), V = cell.a11 cell.a22 cell.a33 (Get_Volume, MD_core/sys.hpp:438-443).
 * All three wait for the PME stream, whose reciprocal virial the caller
 * passes, as the reference does, in d_atom_virial[0]. */

/* system_information::Get_Potential_to_stress (sys.hpp:493-510, kernels :50-85):
 * d_virial_tensor[0] += sum of d_atom_virial[0..atom_numbers); then
 * d_stress[0] += d_virial_tensor[0] / V.  d_virial_tensor may be NULL
 * (the sum alone is added). */
int snb_get_potential_to_stress(snb_context* ctx, int atom_numbers,
                                const snb_ltmatrix3* d_atom_virial, snb_ltmatrix3 cell,
                                snb_ltmatrix3* d_virial_tensor, snb_ltmatrix3* d_stress);

/* system_information::Get_Kinetic_to_stress (sys.hpp:512-524, kernel :14-48):
 * d_stress[0] += sum m v (x) v / V, in the Get_Virial_From_Force_Dis layout
 * (a21 = 2 m vx vy, a31 = 2 m vx vz, a32 = 2 m vy vz). */
int snb_get_kinetic_to_stress(snb_context* ctx, int atom_numbers, const snb_vector* d_vel,
                              const float* d_mass, snb_ltmatrix3 cell, snb_ltmatrix3* d_stress);

/* MD_INFORMATION::Get_pressure (MD_core.cpp:948-975): the two calls above in
 * one pass over the atoms, then, when h_stress or h_pressure is given, a
 * synchronous copy of d_stress[0] and pressure = (a11 + a22 + a33) / 3
 * (internal units; x 6.946827162543585e4 gives bar, common.h:142).  With both
 * NULL nothing is copied and the call can be captured in a graph. */
int snb_get_pressure(snb_context* ctx, int atom_numbers, const snb_vector* d_vel,
                     const float* d_mass, const snb_ltmatrix3* d_atom_virial, snb_ltmatrix3 cell,
                     snb_ltmatrix3* d_virial_tensor, snb_ltmatrix3* d_stress,
                     snb_ltmatrix3* h_stress, float* h_pressure);

/* -------------------------------------------------------------------- PME */
/* Particle_Mesh::Initial (PM_force.cpp:293-652) with the controller keys
 * resolved by the caller: PM_Direct_Tolerance (1e-5), PM_fftx/y/z (<= 0 =
 * automatic from box_length / PM_grid_spacing, common.cpp:639-659),
 * PME_update_interval (1).  beta comes from the tolerance (:244-272).
 * no_direct_interaction_virtual_atom_numbers (:296, :467-479: extra
 * reciprocal-only atoms past atom_numbers) must be 0: any other value returns
 * SNB_ERR_UNSUPPORTED, since this library's atoms are the context's. */
int snb_pme_initial(snb_context* ctx, snb_ltmatrix3 cell, snb_ltmatrix3 rcell,
                    snb_vector box_length, float cutoff,
                    int no_direct_interaction_virtual_atom_numbers, float tolerance,
                    int fftx, int ffty, int fftz, float grid_spacing,
                    int update_interval);
int snb_pme_parameters(snb_context* ctx, float* beta, int* fftx, int* ffty, int* fftz);

/* Particle_Mesh::Update_Box (PM_force.cpp:1251-1263, kernel up_box_bc :1168-1237):
 * the small box change of a barostat step (Main_Box_Change, main.cpp:886-895).
 * Recomputes the influence function (FP64, stored as float) and the
 * neutralising factor for the new cell on the context stream; the grid size,
 * beta and the brick layout stay.  g and dt only move the reference's domain
 * corners and are not used.  A large change re-initialises instead
 * (snb_pme_initial, as Main_Box_Change_Largely does). */
int snb_pme_update_box(snb_context* ctx, snb_ltmatrix3 cell, snb_ltmatrix3 rcell,
                       snb_ltmatrix3 g, float dt);

/* Particle_Mesh::MPI_PME_Excluded_Force_With_Atom_Energy (PM_force.cpp:1816-1839,
 * kernel :1743-1814) on one rank: removes the erf part of the Ewald
 * interaction of every excluded pair.  atom_local/atom_local_id may be NULL;
 * the exclusion arrays may be NULL when snb_set_exclusions was called. */
int snb_pme_excluded_force(snb_context* ctx, int atom_numbers, const int* d_atom_local,
                           const int* d_atom_local_id, const snb_vector* d_crd,
                           snb_ltmatrix3 cell, snb_ltmatrix3 rcell, const float* d_charge,
                           const int* d_excluded_list_start, const int* d_excluded_list,
                           const int* d_excluded_numbers, snb_vector* d_frc,
                           int need_energy, float* d_atom_energy, int need_virial,
                           snb_ltmatrix3* d_atom_virial);

/* Allocation-free variant of snb_pme_excluded_force.  It has no counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc), and it needs no context.  The caller passes inputs and outputs
 * (device pointers, the exclusion CSR of snb_set_exclusions, beta of
 * snb_pme_parameters) and the stream; frc, atom_energy and atom_virial are
 * accumulated (+=) as by snb_pme_excluded_force, bit for bit.  The fixed-point
 * accumulators live in a library-owned device array of SNB_NOALLOC_MAX_ATOMS
 * x 80 bytes (10 MiB) per GPU, held until the process ends; larger systems
 * return SNB_ERR_TOO_LARGE.  Every call on one GPU shares that array and
 * clears what it uses, so calls that overlap on one GPU race; calls must be
 * ordered (one stream, or events).  The first call on a device looks up the
 * array's device address (hipGetSymbolAddress); no call allocates. */
#define SNB_NOALLOC_MAX_ATOMS 131072
int snb_pme_excluded_force_noalloc(int atom_numbers, const snb_vector* d_crd,
                                   snb_ltmatrix3 cell, snb_ltmatrix3 rcell,
                                   const float* d_charge, float beta,
                                   const int* d_excluded_list_start, const int* d_excluded_list,
                                   const int* d_excluded_numbers, snb_vector* d_frc,
                                   int need_energy, float* d_atom_energy, int need_virial,
                                   snb_ltmatrix3* d_atom_virial, void* hip_stream);

/* Particle_Mesh::PME_Reciprocal_Force_With_Energy_And_Virial (PM_force.cpp:1067-1165):
 * B-spline spread, 3-D R2C FFT, influence-function convolution, C2R, force
 * gather; with need_energy adds reciprocal + self + neutralising energy to
 * d_potential[0]; with need_virial adds the reciprocal virial to d_virial[0].
 * rcell must be the one of the last snb_pme_initial / snb_pme_update_box
 * (SNB_ERR_ARGUMENT otherwise: the influence function belongs to a box).
 * With PME_update_interval k > 1 (:1079-1163) the force, scaled by k, and the
 * virial come only on steps with step % k == 0; on the other steps nothing is
 * added to d_frc or d_virial, and the reciprocal energy reported is the one of
 * the last update step's grids (0 before the first), plus the self and
 * neutralising terms of the current charges. */
int snb_pme_reciprocal_force(snb_context* ctx, const snb_vector* d_crd, snb_ltmatrix3 cell,
                             snb_ltmatrix3 rcell, const float* d_charge, snb_vector* d_frc,
                             int need_virial, int need_energy, snb_ltmatrix3* d_virial,
                             float* d_potential, int step);

/* Particle_Mesh::Step_Print (PM_force.cpp:1841-1866): the four PM terms of
 * the last energy evaluation (FP64).  Synchronises. */
int snb_pme_energy(snb_context* ctx, double* direct, double* reciprocal, double* self,
                   double* correction);
/* Reciprocal virial of the last need_virial call (FP64, LTMatrix3 order). */
int snb_pme_reciprocal_virial(snb_context* ctx, double* virial6);

/* ---------------------------------------------------- fixed-point accumulators */
/* d_frc[i] += accumulated force, then clears the accumulator. */
int snb_flush_forces(snb_context* ctx, snb_vector* d_frc);
/* For outputs deferred with NULL: d_atom_energy[i] += the per-atom energies of
 * the last energy calls (LJ + direct Coulomb + exclusion correction + 1-4 +
 * bond) and
 * d_atom_direct_pme_energy[i] = the direct Coulomb part; either may be NULL. */
int snb_flush_atom_energy(snb_context* ctx, float* d_atom_energy, float* d_atom_direct_pme_energy);
/* d_atom_virial[i] += the accumulated per-atom virial, then clears it. */
int snb_flush_atom_virial(snb_context* ctx, snb_ltmatrix3* d_atom_virial);
/* Copies the int64 force accumulators (SoA x[n] | y[n] | z[n]) to d_out. */
int snb_copy_fixed_forces(snb_context* ctx, int64_t* d_out);
int snb_clear_fixed_forces(snb_context* ctx);
/* Copies the int64 per-atom accumulators: which = 0 LJ energy, 1 direct
 * Coulomb energy, 2 exclusion-correction energy, 4 1-4 energy, 5 bond energy
 * (n each), 3 virial (6 x n, SoA). */
int snb_copy_fixed_atom_terms(snb_context* ctx, int which, int64_t* d_out);
/* Copies the PME real-space grids of the last reciprocal call (float,
 * [fftx][ffty][fftz]): which = 0 charge grid Q, 1 potential grid (C2R output). */
int snb_copy_pme_grid(snb_context* ctx, int which, float* d_out);
/* SPONGE_FFT_WRAPPER::R2C / C2R (utils/fft.hpp:93-109) on the PME grid of
 * this context, with the library's own radix kernels: unnormalised 3-D
 * real-to-complex transform of float[fftx][ffty][fftz] into complex
 * float2[fftx][ffty][fftz/2+1] (exponent sign -1), and the inverse (+1). */
int snb_fft_r2c(snb_context* ctx, const float* d_real, float* d_complex);
int snb_fft_c2r(snb_context* ctx, const float* d_complex, float* d_real);

/* ------------------------------------------------ FP64 reference evaluation */
/* No counterpart in the reference (which builds in float): the whole force
 * path of this context in double precision, for precision studies of the
 * FP32 / int64 path (an FP64 trajectory from the same start, compared frame by
 * frame and by energy drift).  Uses the context's modules (LJ tables, 1-4
 * pairs, bonds, exclusions, PME grid and beta, SETTLE triangles) with double
 * coordinates d_crd[3n] and forces d_frc[3n] (accumulated, +=):
 * pairs over the reference-layout half list (the context's list, rebuilt from
 * a float copy of d_crd by the NEIGHBOR_LIST::Update policy at this step; the
 * list export is switched on), each pair, the exclusion correction, the 1-4
 * pairs and the bonds in double; PME reciprocal with a double spread, direct
 * FP64 DFTs along each axis, the influence function on the full grid and a
 * double gather.  Forces are summed in int64 at 2^40, energies and the virial
 * at 2^32 (order-independent).  With need_energy, h_terms[9] receives LJ_short,
 * LJ_long, PM_direct, PM_reciprocal, PM_self, PM_correction, nb14_LJ,
 * nb14_EE, bond; with need_virial, h_virial[6] the total virial tensor of the
 * pair, exclusion, 1-4, bond and reciprocal terms (a11 a21 a22 a31 a32 a33,
 * off-diagonal entries as symmetrised sums); both synchronise the stream. */
int snb_f64_force(snb_context* ctx, const double* d_crd, const float* d_charge,
                  snb_ltmatrix3 cell, snb_ltmatrix3 rcell, int step, int forced_nl,
                  double* d_frc, int need_energy, int need_virial, double* h_terms,
                  double* h_virial);
/* One FP64 NVE step after snb_f64_force (no counterpart in the reference):
 * SETTLE's remember, leap-frog v += f/m dt, x += v dt (NVE_iteration::Leap_Frog,
 * MD_core/nve.hpp:3-48) and SETTLE with v_factor = x_factor = 1 on the
 * context's triangles, all in double (dt in the internal time unit). */
int snb_f64_nve_iterate(snb_context* ctx, const double* d_mass, const double* d_inverse_mass,
                        const double* d_frc, double* d_vel, double* d_crd,
                        snb_ltmatrix3 cell, snb_ltmatrix3 rcell, double dt);

/* ------------------------------------------------------------ profiling */
/* Kernel timing with HIP events recorded on the context stream around each
 * launch of a kernel group (no counterpart in the reference, whose
 * TIME_RECORDER synchronises the device, utils/control/time_recorder.hpp:9-43):
 * SNB_PROF_PAIR = the LJ + direct Coulomb kernel alone, SNB_PROF_NL = the
 * neighbor-list update (check + rebuild kernels), SNB_PROF_PME = the
 * reciprocal pipeline (prep .. gather), SNB_PROF_EXCL = the exclusion
 * correction kernel.  snb_profile_read synchronises and returns the summed
 * milliseconds and launch count since the last reset. */
#define SNB_PROF_PAIR 0
#define SNB_PROF_NL 1
#define SNB_PROF_PME 2
#define SNB_PROF_EXCL 3
int snb_profile_enable(snb_context* ctx, int enable);
int snb_profile_read(snb_context* ctx, int group, int reset, double* total_ms, int* launches);

/* ------------------------------------------------------- step batches */
/* One dispatch per kernel for the same step of several contexts (no
 * counterpart in the reference, which runs one system per process): for a
 * server that advances many small systems in lockstep, where a step of each
 * is a chain of short kernels and the device's dispatch rate, not its compute
 * units, bounds the throughput.
 *
 * Between snb_batch_record_begin(b, ctx) and snb_batch_record_end(b) the
 * entry points called for ctx (on this thread) launch nothing: they record
 * their launches, with the arguments they would pass, after every check and
 * host-side decision they make.  Record each member's calls (for example the
 * steps of one neighbour-list period), then snb_batch_plan merges the
 * records: launches of a kernel that has a batched form become one dispatch
 * over a device table with one entry per member, the others stay as they
 * were, and every member's launches keep their order.  snb_batch_launch
 * issues the plan on one stream, any number of times (it can be captured in a
 * HIP graph); each member's results are bit-identical to making its calls
 * alone.  The recorded arguments are fixed: replaying a plan repeats the
 * recorded steps' host-side decisions (for a fixed refresh interval, record
 * one whole period).  Members need the reciprocal pipeline on the context
 * stream (snb_set_pme_overlap(ctx, 0)) and profiling off
 * (SNB_ERR_UNSUPPORTED otherwise).  A member must not be destroyed, and its
 * caller buffers not freed, while a plan that holds it may still run; the
 * batch itself (its device table) likewise.  Recording is per thread (the
 * calls between begin and end on the recording thread); make plans from one
 * thread at a time.  If a recorded call fails, or snb_batch_plan does,
 * destroy the batch: its records are incomplete.  Set-up calls take effect
 * when made, not when a plan runs: make them outside a recording
 * (snb_nl_set_window_shape and snb_pme_update_box return SNB_ERR_UNSUPPORTED
 * inside one). */
typedef struct snb_batch snb_batch;
int snb_batch_create(snb_batch** out, int device);
int snb_batch_destroy(snb_batch* batch);
int snb_batch_record_begin(snb_batch* batch, snb_context* ctx);
int snb_batch_record_end(snb_batch* batch);
/* launches = dispatches per snb_batch_launch, batched = how many of them are
 * batched forms */
int snb_batch_plan(snb_batch* batch, int* launches, int* batched);
int snb_batch_launch(snb_batch* batch, void* hip_stream);
/* members recorded, dispatches per launch, batched dispatches, and the member
 * launches they replace */
int snb_batch_info(snb_batch* batch, int* members, int* launches, int* batched, int* merged);

#ifdef __cplusplus
}
#endif
#endif /* SPONGE_NB_H */
