/*
 * micm_hip.h -- C-ABI of the MI355X-native Rosenbrock chemistry solver.
 *
 * The boundary replaces the C++-ABI "Driver" seam of the reference's CUDA
 * backend (plain functions taking POD parameter structs, e.g.
 * include/micm/cuda/process/cuda_process_set.cuh:29-41) with extern "C"
 * entry points that take plain pointers and sizes, return an int status and
 * never throw.  Every per-cell array handed to a kernel entry point lives in
 * device memory in the reference's vector ordering with a group size of one
 * wavefront (64 cells):
 *
 *     dense  (cell c, variable v):  ((c / 64) * n_vars + v) * 64 + c % 64
 *     sparse (cell c, non-zero p):  ((c / 64) * nnz    + p) * 64 + c % 64
 *
 * which is include/micm/util/vector_matrix.hpp:200 (VectorMatrix<T, 64>) and
 * include/micm/util/sparse_matrix_vector_ordering_compressed_sparse_row.hpp:79-89
 * (SparseMatrixVectorOrderingCompressedSparseRow<64>).  Buffers hold
 * ceil(n_cells / 64) full groups; padding lanes are never read as results.
 *
 * Species and reactions use the reference's post-Build orders: species are
 * permuted by the diagonal Markowitz reordering
 * (include/micm/solver/solver_builder.inl:78-119) unless the description
 * disables it, and reactions are stable-sorted by rate-constant type
 * (solver_builder.inl:274-329).  The getters below report both permutations.
 */
#ifndef MICM_HIP_H
#define MICM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define MICM_HIP_OK 0
#define MICM_HIP_ERR_INVALID_ARGUMENT (-1)
#define MICM_HIP_ERR_DEVICE (-2)
#define MICM_HIP_ERR_OUT_OF_MEMORY (-3)
#define MICM_HIP_ERR_UNSUPPORTED (-4)
#define MICM_HIP_ERR_SPECIES (-5) /* unknown / unused species, as MicmException in process_set.hpp:291-309 */

/* Cells per vector group (one CDNA wavefront). */
#define MICM_HIP_GROUP 64

/* Doubles per reaction in micm_hip_mechanism_desc::rate_params. */
#define MICM_HIP_RATE_PARAM_STRIDE 24

/* Rate-constant types, numbered in the reference's sort order
 * (solver_builder.inl:279-324).  Parameter layout per reaction row of
 * rate_params (unused slots ignored):
 *   ARRHENIUS    A, B, C, D, E                       (arrhenius_rate_constant.hpp)
 *   TROE         k0_A, k0_B, k0_C, kinf_A, kinf_B, kinf_C, Fc, N
 *   TERNARY      same as TROE                        (ternary_chemical_activation_rate_constant.hpp)
 *   BRANCHED     branch (0 alkoxy, 1 nitrate), X, Y, a0, n
 *   TUNNELING    A, B, C
 *   TAYLOR       A, B, C, D, E, n_coefficients, c_0 .. c_15
 *   REVERSIBLE   A, C, k_r
 *   USER_DEFINED scaling_factor       (one custom parameter per reaction)
 *   SURFACE      diffusion_coefficient, molecular_weight, reaction_probability
 *                                     (two custom parameters per reaction)
 */
#define MICM_HIP_RATE_ARRHENIUS 0
#define MICM_HIP_RATE_TROE 1
#define MICM_HIP_RATE_TERNARY 2
#define MICM_HIP_RATE_BRANCHED 3
#define MICM_HIP_RATE_TUNNELING 4
#define MICM_HIP_RATE_TAYLOR 5
#define MICM_HIP_RATE_REVERSIBLE 6
#define MICM_HIP_RATE_USER_DEFINED 7
#define MICM_HIP_RATE_SURFACE 8
#define MICM_HIP_RATE_N_TYPES 9

/* Solver states, numbered as micm::SolverState (solver_result.hpp:13-35). */
#define MICM_HIP_STATE_NOT_YET_CALLED 0
#define MICM_HIP_STATE_RUNNING 1
#define MICM_HIP_STATE_CONVERGED 2
#define MICM_HIP_STATE_MAX_STEPS 3
#define MICM_HIP_STATE_STEP_SIZE_TOO_SMALL 4
#define MICM_HIP_STATE_REPEATEDLY_SINGULAR 5
#define MICM_HIP_STATE_NAN_DETECTED 6
#define MICM_HIP_STATE_INF_DETECTED 7

/* Per-cell statistics, in the field order of micm::SolverStats
 * (solver_result.hpp:37-57); final_time is reported separately. */
#define MICM_HIP_STAT_FUNCTION_CALLS 0
#define MICM_HIP_STAT_JACOBIAN_UPDATES 1
#define MICM_HIP_STAT_STEPS 2
#define MICM_HIP_STAT_ACCEPTED 3
#define MICM_HIP_STAT_REJECTED 4
#define MICM_HIP_STAT_DECOMPOSITIONS 5
#define MICM_HIP_STAT_SOLVES 6
#define MICM_HIP_N_STATS 7

/* Tableau presets (rosenbrock_solver_parameters.hpp:152-433). */
#define MICM_HIP_ROS_TWO_STAGE 2
#define MICM_HIP_ROS_THREE_STAGE 3
#define MICM_HIP_ROS_FOUR_STAGE 4
#define MICM_HIP_ROS_FOUR_STAGE_DA 41
#define MICM_HIP_ROS_SIX_STAGE_DA 61

/* Integration forms of micm_hip_solve. */
#define MICM_HIP_SOLVE_LOCKSTEP 0   /* one step size per state, as rosenbrock.inl:12-304 */
#define MICM_HIP_SOLVE_PLAIN 1      /* per-cell step size, one launch per phase of an attempt */
#define MICM_HIP_SOLVE_PERSISTENT 2 /* per-cell step size, persistent work-queue kernel; cells that
                                       reach the hand-off step count continue in the cooperative form */
#define MICM_HIP_SOLVE_COOPERATIVE 3 /* per-cell step size, one wavefront per cell, working set in LDS */
/* OR-ed into the form: UpdateStateParameters fused into the Solve launch -- each
 * cell's rate constants are computed from its conditions and custom parameters as
 * the cell is loaded (and stored in the state), instead of by a separate
 * micm_hip_update_state_parameters call.  Same results. */
#define MICM_HIP_SOLVE_UPDATE_RATES 0x100

/* Flat mechanism description.  Species are the phase's non-parameterized
 * species in declaration order; reactions are in the order given to
 * SolverBuilder::SetReactions. */
typedef struct micm_hip_mechanism_desc {
  int64_t n_species;
  int64_t n_reactions;
  const int64_t* reactant_offsets; /* [n_reactions + 1] */
  const int64_t* reactants;        /* species index, or -(p + 1) for parameterized species p */
  const int64_t* product_offsets;  /* [n_reactions + 1] */
  const int64_t* products;         /* species index, or -(p + 1) (skipped, process_set.hpp:301) */
  const double* yields;            /* one per product */
  const int32_t* rate_types;       /* [n_reactions], MICM_HIP_RATE_* */
  const double* rate_params;       /* [n_reactions * MICM_HIP_RATE_PARAM_STRIDE] */
  int64_t n_parameterized;         /* parameterized ("third-body") species */
  const double* parameterized;     /* [n_parameterized * 4]: c0, c_T, c_P, c_rho (parameterized_function.hpp:36) */
  const double* absolute_tolerances; /* [n_species] in declaration order, or NULL for 1e-3 (solver_builder.inl:208) */
  int32_t reorder_state;           /* 1: diagonal Markowitz order (reference default), 0: declaration order */
  int32_t ignore_unused_species;   /* 1: reference default (solver_builder.hpp:64) */
} micm_hip_mechanism_desc;

typedef struct micm_hip_mechanism_info {
  int64_t n_species;
  int64_t n_reactions;
  int64_t n_custom_parameters;
  int64_t nnz_jacobian;   /* including the diagonal (jacobian.hpp:27-30) */
  int64_t nnz_lower;      /* including the unit diagonal */
  int64_t nnz_upper;      /* including the diagonal */
  int64_t n_process_infos;
  int64_t n_jacobian_terms;
  int64_t lu_multiply_adds; /* Doolittle L*U products per factorization */
  int64_t n_reactant_slots;
  int64_t n_product_slots;
} micm_hip_mechanism_info;

/* micm::RosenbrockSolverParameters (rosenbrock_solver_parameters.hpp:22-67). */
typedef struct micm_hip_rosenbrock_parameters {
  int64_t stages;
  int64_t max_number_of_steps;
  double round_off;
  double factor_min;
  double factor_max;
  double rejection_factor_decrease;
  double safety_factor;
  double h_min;
  double h_max;
  double h_start;
  int32_t new_function_evaluation[6];
  double estimator_of_local_order;
  double a[15];
  double c[15];
  double m[6];
  double e[6];
  double alpha[6];
  double gamma[6];
} micm_hip_rosenbrock_parameters;

typedef struct micm_hip_mechanism micm_hip_mechanism;
typedef struct micm_hip_state micm_hip_state;

/* Device addresses of a state's arrays (grouped layout above). */
typedef struct micm_hip_state_pointers {
  double* variables;         /* n_species */
  double* conditions;        /* 3: temperature [K], pressure [Pa], air density [mol m-3] */
  double* custom_parameters; /* n_custom_parameters */
  double* rate_constants;    /* n_reactions */
  double* jacobian;          /* nnz_jacobian (holds -J, as the reference) */
  double* lower;             /* nnz_lower */
  double* upper;             /* nnz_upper */
  double* absolute_tolerances; /* [n_species], not grouped */
  int32_t* cell_state;       /* [n_cells] MICM_HIP_STATE_* of the last solve */
  int64_t* cell_stats;       /* [n_cells * MICM_HIP_N_STATS] */
  double* cell_final_time;   /* [n_cells] */
  int64_t n_cells;
  int64_t n_groups;
} micm_hip_state_pointers;

/* ------------------------------------------------------------ errors */
/* Message of the last failing call on this thread (never NULL). */
const char* micm_hip_last_error(void);

/* ------------------------------------------------------------ solver parameters */
/* RosenbrockSolverParameters::*StageRosenbrockParameters (rosenbrock_solver_parameters.hpp:152-433). */
int micm_hip_rosenbrock_parameters_preset(int32_t preset, micm_hip_rosenbrock_parameters* out);

/* ------------------------------------------------------------ mechanism */
/* Replaces SolverBuilder::Build's symbolic work (solver_builder.inl:227-548):
 * reaction sort, species map, ProcessSet tables (process_set.hpp:256-492),
 * Jacobian pattern (jacobian.hpp:15-32), Doolittle fill and schedule
 * (lu_decomposition_doolittle.inl:69-299), LinearSolver tables
 * (linear_solver.inl:166-224) and the rate-constant store
 * (reaction_rate_store.hpp:221-368); uploads every table to the current device. */
int micm_hip_mechanism_create(const micm_hip_mechanism_desc* desc, micm_hip_mechanism** out);
void micm_hip_mechanism_destroy(micm_hip_mechanism* mechanism);
int micm_hip_mechanism_get_info(const micm_hip_mechanism* mechanism, micm_hip_mechanism_info* out);
/* state_to_declared[i] = declared species index held in state column i. */
int micm_hip_mechanism_get_species_order(const micm_hip_mechanism* mechanism, int64_t* state_to_declared);
/* sorted_to_declared[r] = declared reaction index of rate-constant column r. */
int micm_hip_mechanism_get_reaction_order(const micm_hip_mechanism* mechanism, int64_t* sorted_to_declared);
/* custom_to_reaction[p] = sorted reaction index whose rate uses custom parameter p. */
int micm_hip_mechanism_get_custom_parameter_owner(const micm_hip_mechanism* mechanism, int64_t* custom_to_reaction);
/* (row, column) of every stored element in storage order; any pointer may be NULL. */
int micm_hip_mechanism_get_patterns(const micm_hip_mechanism* mechanism,
                                    int64_t* jacobian_rows, int64_t* jacobian_cols,
                                    int64_t* lower_rows, int64_t* lower_cols,
                                    int64_t* upper_rows, int64_t* upper_cols);
/* Storage index of each diagonal element of the Jacobian (state.inl:313). */
int micm_hip_mechanism_get_jacobian_diagonal(const micm_hip_mechanism* mechanism, int64_t* diagonal);

/* ------------------------------------------------------------ kernel entry points
 * All pointers are device pointers in the grouped layout; `stream` is a
 * hipStream_t (NULL = the legacy default stream).  Launches are asynchronous. */

/* CalculateRateConstantsKernelDriver (cuda_rate_constant_kernel.cuh:68-73) and the
 * parameterized multipliers (reaction_rate_store.hpp:510-517): conditions hold
 * 3 values per cell (T, P, air density). */
int micm_hip_calculate_rate_constants(const micm_hip_mechanism* mechanism, int64_t n_cells,
                                      const double* conditions, const double* custom_parameters,
                                      double* rate_constants, void* stream);
/* AddForcingTermsKernelDriver (cuda_process_set.cuh:29-35): forcing += f(y). */
int micm_hip_add_forcing_terms(const micm_hip_mechanism* mechanism, int64_t n_cells,
                               const double* rate_constants, const double* variables,
                               double* forcing, void* stream);
/* The pair Rosenbrock::Solve issues at each step start, initial_forcing.Fill(0) then
 * AddForcingTerms (rosenbrock.inl:98-99), in one pass: forcing = f(y), bit for bit what the
 * pair gives; forcing is only written. */
int micm_hip_fill_and_add_forcing_terms(const micm_hip_mechanism* mechanism, int64_t n_cells,
                                        const double* rate_constants, const double* variables,
                                        double* forcing, void* stream);
/* SubtractJacobianTermsKernelDriver (cuda_process_set.cuh:37-41): jacobian -= J(y). */
int micm_hip_subtract_jacobian_terms(const micm_hip_mechanism* mechanism, int64_t n_cells,
                                     const double* rate_constants, const double* variables,
                                     double* jacobian, void* stream);
/* The pair jacobian_.Fill(0) then SubtractJacobianTerms (rosenbrock.inl:109-110) in one
 * pass: jacobian = -J(y), bit for bit what the pair gives; jacobian is only written. */
int micm_hip_fill_and_subtract_jacobian_terms(const micm_hip_mechanism* mechanism, int64_t n_cells,
                                              const double* rate_constants, const double* variables,
                                              double* jacobian, void* stream);
/* AlphaMinusJacobianDriver (cuda_rosenbrock.cuh:18-22): jacobian[diag] += alpha. */
int micm_hip_alpha_minus_jacobian(const micm_hip_mechanism* mechanism, int64_t n_cells,
                                  double* jacobian, double alpha, void* stream);
/* LuDecompositionDoolittle::Decompose (lu_decomposition_doolittle.inl:325-403). */
int micm_hip_lu_decompose(const micm_hip_mechanism* mechanism, int64_t n_cells,
                          const double* jacobian, double* lower, double* upper, void* stream);
/* LinearSolver::Solve (linear_solver.inl:235-294): x <- U^-1 L^-1 x in place. */
int micm_hip_lu_solve(const micm_hip_mechanism* mechanism, int64_t n_cells,
                      const double* lower, const double* upper, double* x, void* stream);
/* Per-cell form of NormalizedError (rosenbrock.inl:353-402):
 * error[c] = max(sqrt(sum_i (e_i / (atol_i + rtol * max(|y_i|, |ynew_i|)))^2 / N), 1e-10).
 * absolute_tolerances is a plain [n_species] device array; error is [n_cells]. */
int micm_hip_normalized_error(const micm_hip_mechanism* mechanism, int64_t n_cells,
                              const double* variables, const double* new_variables,
                              const double* errors, const double* absolute_tolerances,
                              double relative_tolerance, double* error, void* stream);

/* NormalizedErrorDriver (cuda_rosenbrock.cuh:29-35) / NormalizedError (rosenbrock.inl:353-402) for a
 * whole State: *error = max(sqrt(sum / (n_cells * n_species)), 1e-10), the sum added one term at a
 * time in the order the reference's dense matrix policy visits the terms -- vector_length 1:
 * Matrix<Real> (cell by cell, species inner); L > 1: VectorMatrix<L> (each group of L cells species
 * by species with the group's cells inner, a partial last group over its own cells;
 * vector_matrix.hpp:860-867).  Arrays grouped as above (the device layout is always 64-wide; the
 * order is the reference State's); absolute_tolerances a plain [n_species] device array; error one
 * device double. */
int micm_hip_normalized_error_state(const micm_hip_mechanism* mechanism, int64_t n_cells,
                                    const double* variables, const double* new_variables, const double* errors,
                                    const double* absolute_tolerances, double relative_tolerance,
                                    int64_t vector_length, double* error, void* stream);

/* ------------------------------------------------------------ state */
/* Solver::GetState (solver.hpp:174-195): device arrays for n_cells plus
 * integrator scratch; tolerances start at the mechanism's defaults with
 * rtol 1e-6 (state.hpp:42). */
int micm_hip_state_create(const micm_hip_mechanism* mechanism, int64_t n_cells, micm_hip_state** out);
void micm_hip_state_destroy(micm_hip_state* state);
int micm_hip_state_get_pointers(const micm_hip_state* state, micm_hip_state_pointers* out);
/* Host transfers use cell-major host arrays [n_cells][n_values] (state order). */
int micm_hip_state_set_variables(micm_hip_state* state, const double* host);
int micm_hip_state_get_variables(const micm_hip_state* state, double* host);
int micm_hip_state_set_conditions(micm_hip_state* state, const double* host); /* [n_cells][3] */
int micm_hip_state_set_custom_parameters(micm_hip_state* state, const double* host);
int micm_hip_state_get_rate_constants(const micm_hip_state* state, double* host);
int micm_hip_state_set_rate_constants(micm_hip_state* state, const double* host);
int micm_hip_state_set_tolerances(micm_hip_state* state, double relative_tolerance,
                                  const double* absolute_tolerances /* [n_species] or NULL */);
/* Per-cell result of the last solve: states [n_cells], stats [n_cells][7], final_time [n_cells]; any may be NULL. */
int micm_hip_state_get_results(const micm_hip_state* state, int32_t* states, int64_t* stats, double* final_time);
/* Summation order of the MICM_HIP_SOLVE_LOCKSTEP form's error norm: that of a reference State over
 * VectorMatrix<vector_length> (1, the default: Matrix<Real>); see micm_hip_normalized_error_state. */
int micm_hip_state_set_lockstep_order(micm_hip_state* state, int64_t vector_length);
/* Stream used by the state's solves (hipStream_t; NULL = default stream). */
int micm_hip_state_set_stream(micm_hip_state* state, void* stream);
/* Step count at which the persistent form hands a cell to the cooperative form
 * (0: never; negative: the library's default).  Results do not depend on it. */
int micm_hip_state_set_handoff(micm_hip_state* state, int64_t steps);
/* Launch shape of the cooperative form: threads per cell (64, 128, 256, 512 or 1024; 0: the
 * library's choice) and which constant tables are staged in LDS (bit 0 the
 * substitution tables, bit 1 the LU tables; -1: the library's choice; tables that
 * do not fit next to the cell's data stay in global memory), or 4: the compact
 * per-cell layout (no tables staged, fewer arrays in LDS so that more cells fit a CU;
 * needs more than 64 threads and at least one thread per species, else ignored), its
 * register budget chosen from occupancy; 5 / 6 / 7: the compact layout with 256 threads built
 * for three waves per SIMD (168 registers) / two (256) / no register bound.  A tuning
 * knob in the role of the vector length L of the reference's GPU policies
 * (cuda_solver_builder.hpp:24-38); results do not depend on it. */
int micm_hip_state_set_cooperative_shape(micm_hip_state* state, int32_t threads, int32_t stage);
/* Blocking of the cooperative form's LU level schedule: the shortest run of products an
 * element's ordered sum (lu_decomposition_doolittle.inl:325-403) is split into, except its
 * last -- 1, 2, 4 or 8; -1 keeps every sum whole; 0: the library's default (4).  Every
 * blocking subtracts the same products in the same order; it only moves where a sum is
 * split between the levels of the schedule.  A tuning knob (the autotuner's "LU schedule
 * blocking"); results do not depend on it. */
int micm_hip_state_set_lu_blocking(micm_hip_state* state, int32_t min_segment);
/* Whether the cooperative form fits this mechanism on the current device (its
 * per-cell working set must fit one workgroup's LDS); *out = 1 or 0. */
int micm_hip_mechanism_cooperative_supported(const micm_hip_mechanism* mechanism, int32_t stages, int32_t* out);

/* Solver::UpdateStateParameters (solver.hpp:217-247) on the device. */
int micm_hip_update_state_parameters(const micm_hip_mechanism* mechanism, micm_hip_state* state);

/* Solver::Solve (solver.hpp:133-142): integrate every cell over time_step with
 * the given form, then clamp negatives to zero (solver.hpp:275-297).  Blocks
 * until the device has finished; per-cell outcomes via micm_hip_state_get_results.
 * The returned status only reports API/device failures. */
int micm_hip_solve(const micm_hip_mechanism* mechanism, micm_hip_state* state, double time_step,
                   const micm_hip_rosenbrock_parameters* parameters, int32_t form);

/* Solver::Solve in the cooperative form (MICM_HIP_SOLVE_COOPERATIVE, flags 0 or
 * MICM_HIP_SOLVE_UPDATE_RATES) without allocating: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under stream capture, tight serving loops).  The
 * launch's scratch (queue counter, the compact layout's rate-constant rows) lives in one static
 * 4 MiB device array per GPU; a launch shape needing more is refused with
 * MICM_HIP_ERR_INVALID_ARGUMENT.  Enqueued on the state's stream without waiting; results are
 * those of micm_hip_solve once the stream has finished.  Every call on one GPU shares that
 * array: calls that overlap on a GPU (other threads or streams) race. */
int micm_hip_solve_noalloc(const micm_hip_mechanism* mechanism, micm_hip_state* state, double time_step,
                           const micm_hip_rosenbrock_parameters* parameters, int32_t flags);

/* ---------------------------------------------------------------- device-to-device transfers
 * Grouped device arrays of n_groups * 64 rows, on the state's stream; the call waits for the copy.
 * CudaDenseMatrix::CopyToDevice / CopyToHost (cuda_dense_matrix.hpp:46-50) for callers whose
 * data already lives on the device (the autograd binding). */
int micm_hip_state_set_variables_device(micm_hip_state* state, const double* grouped);
int micm_hip_state_get_variables_device(const micm_hip_state* state, double* grouped);
int micm_hip_state_set_rate_constants_device(micm_hip_state* state, const double* grouped);
int micm_hip_state_get_rate_constants_device(const micm_hip_state* state, double* grouped);

/* ---------------------------------------------------------------- discrete adjoint
 * No counterpart in the reference, which has no derivatives of Solve.  What is
 * differentiated is the map a recorded Solve applied to each cell: its accepted steps
 * (rosenbrock.inl:140-207, one Jacobian and factorisation per step), each with its step
 * size H frozen, and the clamp of solver.hpp:275-297 after every Solve.  Rejected attempts
 * only chose H and drop out.
 *
 * A tape holds per cell up to max_records checkpoints of n_species + 1 doubles:
 * (H, y at the start of an accepted step) or (-1, y before the post-solve clamp).  The
 * adjoint replays each step from its checkpoint (stage vectors recomputed, not stored),
 * then runs its stages backwards with the transposed factors. */
typedef struct micm_hip_tape micm_hip_tape;
int micm_hip_tape_create(const micm_hip_mechanism* mechanism, int64_t n_cells, int64_t max_records,
                         micm_hip_tape** out);
void micm_hip_tape_destroy(micm_hip_tape* tape);
/* Forget every record: the next recorded Solve starts a new sequence. */
int micm_hip_tape_reset(micm_hip_tape* tape);
/* Host copies of cells [first_cell, first_cell + n_cells): records [n_cells][max_records][n_species + 1],
 * counts [n_cells]; either may be NULL.  A count above max_records marks a cell whose tape
 * overflowed (its later records were dropped). */
int micm_hip_tape_get(const micm_hip_tape* tape, int64_t first_cell, int64_t n_cells, double* records,
                      int64_t* counts);
int micm_hip_tape_set(micm_hip_tape* tape, const double* records, const int64_t* counts);
int micm_hip_tape_overflowed(const micm_hip_tape* tape, int64_t* n_cells_overflowed);
/* Solver::Solve (solver.hpp:133-142) in the cooperative form, appending each cell's
 * checkpoints to the tape (results identical to micm_hip_solve).  flags: 0 or
 * MICM_HIP_SOLVE_UPDATE_RATES.  Successive recorded Solves extend the same tape. */
int micm_hip_solve_recorded(const micm_hip_mechanism* mechanism, micm_hip_state* state, double time_step,
                            const micm_hip_rosenbrock_parameters* parameters, int32_t flags, micm_hip_tape* tape);
/* Reverse sweep over the tape with the state's rate constants (those of the recorded
 * Solves).  Device arrays, grouped: seed = dL/dy_final [n_species], grad_variables =
 * dL/dy_initial [n_species], grad_log_rate_constants = k dL/dk [n_reactions] per cell (for
 * every rate type the derivative with respect to the log of a common factor on the
 * reaction's A-prefactors; sum over cells for a shared parameter).  Cells whose tape
 * overflowed get NaN.  The parameters must be those of the recorded Solves. */
int micm_hip_adjoint(const micm_hip_mechanism* mechanism, micm_hip_state* state,
                     const micm_hip_rosenbrock_parameters* parameters, const micm_hip_tape* tape,
                     const double* seed, double* grad_variables, double* grad_log_rate_constants);

/* ---------------------------------------------------------------- packed solves
 * Several states -- of one mechanism or of different ones -- solved together.  The
 * per-cell Solve of every state (MICM_HIP_SOLVE_PERSISTENT or _COOPERATIVE) is
 * captured once into a hipGraph, one state after another (each launch fills the
 * GPU; run as concurrent branches, launches of different mechanisms fragment the
 * CUs' LDS and finish later), and later solves with the same time step, parameters
 * and launch shapes replay it.  The
 * states must outlive the packed object and must not be used by another call while
 * a packed solve runs.  No counterpart in the reference, whose States are solved one
 * at a time; for callers that integrate several mechanisms (or several cell sets)
 * per time step. */
typedef struct micm_hip_packed micm_hip_packed;
int micm_hip_packed_create(int32_t n_states, micm_hip_state* const* states, int32_t form, micm_hip_packed** out);
/* Blocks until every state has finished; per-cell results via micm_hip_state_get_results. */
int micm_hip_packed_solve(micm_hip_packed* packed, double time_step, const micm_hip_rosenbrock_parameters* params);
/* Capture and launch the packed solves on `stream` (a hipStream_t other than the legacy default
 * stream, which cannot be captured) instead of the object's own stream. */
int micm_hip_packed_set_stream(micm_hip_packed* packed, void* stream);
/* 1 if the last micm_hip_packed_solve replayed a graph captured earlier, else 0. */
int micm_hip_packed_replayed(const micm_hip_packed* packed);
void micm_hip_packed_destroy(micm_hip_packed* packed);

#ifdef __cplusplus
}
#endif

#endif /* MICM_HIP_H */
