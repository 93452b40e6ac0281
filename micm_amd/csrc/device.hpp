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
    const int32_t* lu_term_l;
    const int32_t* lu_term_u;
    const int32_t* lower_diag;
    const int32_t* upper_diag;
    // substitution
    const int32_t* fwd_offsets;
    const int32_t* fwd_mat;
    const int32_t* fwd_col;
    const int32_t* bwd_offsets;
    const int32_t* bwd_mat;
    const int32_t* bwd_col;
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
  };

  // Per-cell integrator state (rosenbrock.inl:75-78 plus SolverStats).
  struct CellControl
  {
    double H, t, last_alpha;
    int32_t flags;   // bit 0 reject_last, bit 1 reject_more, bit 2 start of a new step
    int32_t status;  // MICM_HIP_STATE_*
    int64_t stats[7];
  };

  constexpr int32_t kRejectLast = 1, kRejectMore = 2, kNewStep = 4;
}  // namespace micm_hip
