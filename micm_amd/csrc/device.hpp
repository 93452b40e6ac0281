// device.hpp -- device-side view of a compiled mechanism and the per-cell
// building blocks of the Rosenbrock path.  One lane of a wavefront owns one
// grid cell; per-cell arrays use the grouped layout of include/micm_hip.h,
// so lane l of a wave touching value v reads base[(group*n_values + v)*64 + l]
// and the 64 lanes together read one contiguous 512-byte segment.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace micm_hip
{
  constexpr int kGroup = 64;

  // Device pointers of one mechanism's tables (all uniform across a wavefront,
  // so the compiler serves them with scalar loads).
  struct DevMech
  {
    int32_t n_species, n_reactions, n_custom;
    int32_t nnz_jac, nnz_lower, nnz_upper;
    int32_t n_terms, n_lu_elements;
    // rate constants
    const int32_t* rate_type;
    const double* rate_params;
    const int32_t* custom_index;
    const int32_t* multiplier_offsets;
    const double* multipliers;
    // forcing
    const int32_t* reactant_offsets;
    const int32_t* reactant_ids;
    const int32_t* product_offsets;
    const int32_t* product_ids;
    const double* yields;
    // Jacobian terms
    const int32_t* term_process;
    const int32_t* term_dep_offsets;
    const int32_t* term_dep_ids;
    const int32_t* term_flat_offsets;
    const int32_t* term_flat;
    const int32_t* jac_diag;
    // Doolittle schedule
    const int32_t* lu_out;
    const int32_t* lu_a;
    const int32_t* lu_div;
    const int32_t* lu_term_offsets;
    const int2* lu_terms;   // (L index, U index)
    const int32_t* lower_diag;
    const int32_t* upper_diag;
    // substitution
    const int32_t* fwd_offsets;
    const int2* fwd_terms;  // (L index, column)
    const int32_t* bwd_offsets;
    const int2* bwd_terms;  // (U index, column)
  };

  // One level schedule (tasks of a level are independent).  tasks[t].x = out | diag << 16
  // (diag 0xffff: none), tasks[t].y = first term; tasks[n_tasks] is a sentinel whose .y
  // closes the last term range.
  struct DevLevels
  {
    const int32_t* offsets;  // n_levels + 1
    const uint2* tasks;      // n_tasks + 1
    const uint32_t* terms;   // two 16-bit positions
    const uint4* records;    // 4 per task + 4 for the sentinel (LevelSchedule::records)
    int32_t n_levels, n_tasks, n_terms;
    int32_t max_level_terms;  // largest number of terms of one level
  };

  // The forward substitution's column sweep (ColumnSweep): its levels from first_level on
  // run as n_cols column steps of one wavefront; lanes[r] = row | start << 16, the row's
  // column mask (low, high word), and the source of column r.
  struct DevSweep
  {
    const uint4* lanes;  // kSweepLanes
    int32_t first_level;  // -1: none
    int32_t n_cols;
  };

  // The back substitution's task chain (TaskChain): its first `levels` levels as n_tasks
  // tasks run in order by one wavefront.
  struct DevChain
  {
    const uint32_t* headers;     // 2 per task, kChainTasks tasks
    const uint32_t* lane_terms;  // kChainLanes per task
    int32_t levels;              // 0: none
    int32_t n_tasks;
  };

  // Tables of the cooperative (one workgroup per cell) form; see MechanismTables.
  struct DevCoop
  {
    int32_t nnz_alu;
    int32_t n_terms;
    const int32_t* jac_to_alu;
    const int32_t* alu_diag;
    const int32_t* alu_lower;  // L storage index -> ALU position (-1: the unit diagonal)
    const int32_t* alu_upper;  // U storage index -> ALU position
    DevLevels lu, fwd, bwd;
    DevSweep fwd_sweep;
    DevChain bwd_chain;
    const uint64_t* rxn_pack;   // reactant ids of each reaction (see MechanismTables)
    const uint64_t* term_pack;  // process and dependent ids of each Jacobian term
    const int32_t* spec_offsets;
    const int32_t* spec_rxn;
    const double* spec_coef;
    const int32_t* jel_offsets;
    const int32_t* jel_term;
    const double* jel_coef;
    const uint64_t* jel_pack;   // term_pack of each Jacobian gather entry
  };

  // One lane's view of a grouped array: element v of this lane's cell.
  struct Lane
  {
    double* p;
    __device__ __forceinline__ double& operator[](int v) const
    {
      return p[static_cast<size_t>(v) * kGroup];
    }
  };

  __device__ __forceinline__ Lane lane_view(double* base, size_t group, int n_values, int lane)
  {
    return Lane{ base + group * static_cast<size_t>(n_values) * kGroup + lane };
  }

  // Tableau and step-control constants of micm::RosenbrockSolverParameters, with
  // the derived h_min / h_max / h_start of rosenbrock.inl:33-36.
  struct DevParams
  {
    int32_t stages;
    int64_t max_steps;
    double round_off, factor_min, factor_max, reject_decrease, safety;
    double h_min, h_max, h_start;
    double inv_elo;  // 1 / estimator_of_local_order
    double gamma0;
    double dt;
    double rtol;
    int32_t new_f[6];
    double a[15], c[15], m[6], e[6];
    int64_t handoff_steps;  // per-lane form hands a cell to the cooperative form at this step count (0: never)
    int64_t error_group;    // lockstep error norm: vector length of the summation order (1: Matrix<Real>)
  };

  // Per-cell integrator state (rosenbrock.inl:75-78 plus SolverStats).
  struct CellControl
  {
    double H, t, last_alpha;
    int32_t flags;   // bit 0 reject_last, bit 1 reject_more, bit 2 start of a new step,
                     // bit 3 Jacobian due (set and cleared within one attempt)
    int32_t status;  // MICM_HIP_STATE_*
    int64_t stats[7];
  };

  constexpr int32_t kRejectLast = 1, kRejectMore = 2, kNewStep = 4, kJacobianDue = 8;

  // Status of a cell handed from the per-lane form to the cooperative form at a step
  // boundary; never visible to callers.
  constexpr int32_t kHandedOff = 100;
}  // namespace micm_hip
