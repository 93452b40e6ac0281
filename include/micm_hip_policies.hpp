// micm_hip_policies.hpp -- the MICM policy set that runs Solver::Solve on the GPU through
// the C-ABI of micm_hip.h.  A header for MICM's side of the boundary: it is compiled against
// the reference's headers (include/micm) and links libmicm_hip.so.
//
//   auto solver = micm::HipSolverBuilder<micm::HipRosenbrockSolverParameters>(
//                     micm::RosenbrockSolverParameters::ThreeStageRosenbrockParameters())
//                     .SetSystem(system).SetReactions(reactions).Build();
//   auto state = solver.GetState(n_cells);
//   ... set conditions and concentrations as for CpuSolverBuilder ...
//   solver.UpdateStateParameters(state);   // rate constants on the host (the reference's store)
//   auto result = solver.Solve(dt, state);  // the whole Rosenbrock Solve on the GPU
//
// It plugs into SolverBuilder's seven policies (solver_builder.hpp:37-44) the way the CUDA
// backend's CudaSolverBuilderInPlace does (cuda/solver/cuda_solver_builder.hpp:24-38):
//
//   SolverParametersPolicy  HipRosenbrockSolverParameters (SolverType = HipRosenbrockSolver,
//                           as cuda_solver_parameters.hpp:13-24)
//   DenseMatrixPolicy       VectorMatrix<Real, L>            (host mirror, vector_matrix.hpp)
//   SparseMatrixPolicy      SparseMatrix<Real, SparseMatrixVectorOrdering<L>>
//   RatesPolicy             HipProcessSet: the reference's ProcessSet, plus the same processes
//                           compiled into a device mechanism (micm_hip_mechanism_create)
//   LuDecompositionPolicy,  the reference's host types; the builder constructs them from the
//   LinearSolverPolicy      Jacobian pattern, the device path does not call them
//   StatePolicy             HipState: the reference's State plus a device state
//                           (micm_hip_state), synced per Solve as CudaState::SyncInputsToDevice /
//                           SyncOutputsToHost (cuda/solver/cuda_state.hpp:40-148)
//
// Semantics: HipRosenbrockSolverParameters::form_ = MICM_HIP_SOLVE_LOCKSTEP (the default) is the
// reference's multi-cell State -- one step size for all cells, the error norm summed in
// VectorMatrix<L> order (rosenbrock.inl:353-402, vector_matrix.hpp:860-867) -- and gives the
// reference CPU solver's results on the same policies bit for bit.  The per-cell forms
// (MICM_HIP_SOLVE_PERSISTENT, MICM_HIP_SOLVE_COOPERATIVE) give every cell its own step size, as
// one 1-cell State per cell; their SolverResult reports the first cell that did not converge
// (or cell 0) and its statistics.  Rate constants stay on the host: Solver::UpdateStateParameters
// computes them with the reference's ReactionRateConstantStore and each Solve uploads them.
// Constraints (DAE) are not supported, as with the CUDA policies.
#pragma once

#include "micm_hip.h"

#include <micm/process/process_set.hpp>
#include <micm/solver/linear_solver.hpp>
#include <micm/solver/lu_decomposition.hpp>
#include <micm/solver/rosenbrock_solver_parameters.hpp>
#include <micm/solver/solver_builder.hpp>
#include <micm/solver/solver_result.hpp>
#include <micm/solver/state.hpp>
#include <micm/util/micm_exception.hpp>
#include <micm/util/sparse_matrix.hpp>
#include <micm/util/sparse_matrix_vector_ordering.hpp>
#include <micm/util/vector_matrix.hpp>

#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace micm
{
  namespace hip_detail
  {
    // A non-zero status of the library as the CUDA backend reports a failed CUDA call
    // (src/util/cuda_util.cu:9-16): a MicmException of the internal-error category.
    inline void check(int rc, const char* what)
    {
      if (rc != MICM_HIP_OK)
        throw MicmException(
            MICM_ERROR_CATEGORY_INTERNAL, MICM_INTERNAL_ERROR_CODE_GENERAL, std::string(micm_hip_last_error()) + " : " + what);
    }

    // The processes in the order the builder sorted them, species by the builder's state
    // index: the device mechanism then uses the reference's own species and reaction orders
    // (reordering off; every reaction of one type, so no second sort).  Rate-constant
    // parameters are not needed on the device: the host computes the rate constants.
    inline std::shared_ptr<micm_hip_mechanism> CompileProcesses(const std::vector<Process>& processes,
                                                                const std::unordered_map<std::string, Index>& species)
    {
      std::vector<int64_t> r_off{ 0 }, reactants, p_off{ 0 }, products;
      std::vector<double> yields, params;
      std::vector<int32_t> types;
      for (const auto& process : processes)
      {
        for (const auto& reactant : process.process_.reactants_)
          if (!reactant.IsParameterized())
            reactants.push_back(static_cast<int64_t>(species.at(reactant.name_)));
        for (const auto& product : process.process_.products_)
          if (!product.species_.IsParameterized())
          {
            products.push_back(static_cast<int64_t>(species.at(product.species_.name_)));
            yields.push_back(product.coefficient_);
          }
        r_off.push_back(static_cast<int64_t>(reactants.size()));
        p_off.push_back(static_cast<int64_t>(products.size()));
        types.push_back(MICM_HIP_RATE_ARRHENIUS);
        params.insert(params.end(), MICM_HIP_RATE_PARAM_STRIDE, 0.0);
      }
      micm_hip_mechanism_desc d{};
      d.n_species = static_cast<int64_t>(species.size());
      d.n_reactions = static_cast<int64_t>(processes.size());
      d.reactant_offsets = r_off.data();
      d.reactants = reactants.data();
      d.product_offsets = p_off.data();
      d.products = products.data();
      d.yields = yields.data();
      d.rate_types = types.data();
      d.rate_params = params.data();
      d.reorder_state = 0;
      d.ignore_unused_species = 1;
      micm_hip_mechanism* m = nullptr;
      check(micm_hip_mechanism_create(&d, &m), "micm_hip_mechanism_create");
      return std::shared_ptr<micm_hip_mechanism>(m, micm_hip_mechanism_destroy);
    }

    inline micm_hip_rosenbrock_parameters ToDevice(const RosenbrockSolverParameters& p)
    {
      micm_hip_rosenbrock_parameters d{};
      d.stages = static_cast<int64_t>(p.stages_);
      d.max_number_of_steps = static_cast<int64_t>(p.max_number_of_steps_);
      d.round_off = p.round_off_;
      d.factor_min = p.factor_min_;
      d.factor_max = p.factor_max_;
      d.rejection_factor_decrease = p.rejection_factor_decrease_;
      d.safety_factor = p.safety_factor_;
      d.h_min = p.h_min_;
      d.h_max = p.h_max_;
      d.h_start = p.h_start_;
      d.estimator_of_local_order = p.estimator_of_local_order_;
      for (int i = 0; i < 6; ++i)
      {
        d.new_function_evaluation[i] = p.new_function_evaluation_[i] ? 1 : 0;
        d.m[i] = p.m_[i];
        d.e[i] = p.e_[i];
        d.alpha[i] = p.alpha_[i];
        d.gamma[i] = p.gamma_[i];
      }
      for (int i = 0; i < 15; ++i)
      {
        d.a[i] = p.a_[i];
        d.c[i] = p.c_[i];
      }
      return d;
    }
  }  // namespace hip_detail

  /// RatesPolicy: the reference's ProcessSet (host tables, Jacobian pattern) and the same
  /// processes compiled for the device.
  template<class DenseMatrixPolicy, class SparseMatrixPolicy>
  class HipProcessSet : public ProcessSet<DenseMatrixPolicy, SparseMatrixPolicy>
  {
   public:
    std::shared_ptr<micm_hip_mechanism> device_;

    HipProcessSet() = default;

    HipProcessSet(const std::vector<Process>& processes, const std::unordered_map<std::string, Index>& variable_map)
        : ProcessSet<DenseMatrixPolicy, SparseMatrixPolicy>(processes, variable_map),
          device_(hip_detail::CompileProcesses(processes, variable_map))
    {
    }

    HipProcessSet(
        const std::vector<Process>& processes,
        const std::unordered_map<std::string, Index>& variable_map,
        const std::vector<ExternalModelProcessSet<DenseMatrixPolicy, SparseMatrixPolicy>>& external_process_sets)
        : ProcessSet<DenseMatrixPolicy, SparseMatrixPolicy>(processes, variable_map, external_process_sets),
          device_(hip_detail::CompileProcesses(processes, variable_map))
    {
      if (!external_process_sets.empty())
        throw MicmException(MICM_ERROR_CATEGORY_INTERNAL, MICM_INTERNAL_ERROR_CODE_GENERAL,
                            "HipProcessSet: external model process sets are not supported on the device");
    }
  };

  /// StatePolicy: the reference's State and its device counterpart, created on the first
  /// Solve with the solver's device mechanism.
  template<class DenseMatrixPolicy, class SparseMatrixPolicy, class LuDecompositionPolicy>
  class HipState : public State<DenseMatrixPolicy, SparseMatrixPolicy, LuDecompositionPolicy>
  {
    using Base = State<DenseMatrixPolicy, SparseMatrixPolicy, LuDecompositionPolicy>;
    std::shared_ptr<micm_hip_state> device_;
    const micm_hip_mechanism* mechanism_ = nullptr;
    std::vector<double> buffer_;

   public:
    HipState() = default;
    HipState(const StateParameters& parameters, const Index number_of_grid_cells)
        : Base(parameters, number_of_grid_cells)
    {
    }
    HipState(HipState&&) noexcept = default;
    HipState& operator=(HipState&&) noexcept = default;

    micm_hip_state* Device() const
    {
      return device_.get();
    }

    /// Concentrations, rate constants and tolerances to the device (the host arrays are
    /// the reference State's own).
    void SyncInputsToDevice(const micm_hip_mechanism* mechanism, int64_t lockstep_order)
    {
      const Index cells = this->number_of_grid_cells_;
      if (!device_ || mechanism_ != mechanism)
      {
        micm_hip_state* s = nullptr;
        hip_detail::check(micm_hip_state_create(mechanism, static_cast<int64_t>(cells), &s), "micm_hip_state_create");
        device_ = std::shared_ptr<micm_hip_state>(s, micm_hip_state_destroy);
        mechanism_ = mechanism;
      }
      auto upload = [&](const DenseMatrixPolicy& m, int (*set)(micm_hip_state*, const double*), const char* what)
      {
        const Index cols = m.NumColumns();
        buffer_.resize(cells * cols);
        for (Index c = 0; c < cells; ++c)
          for (Index i = 0; i < cols; ++i)
            buffer_[c * cols + i] = m[c][i];
        hip_detail::check(set(device_.get(), buffer_.data()), what);
      };
      upload(this->variables_, micm_hip_state_set_variables, "micm_hip_state_set_variables");
      upload(this->rate_constants_, micm_hip_state_set_rate_constants, "micm_hip_state_set_rate_constants");
      std::vector<double> atol(this->absolute_tolerance_.size());
      for (std::size_t i = 0; i < atol.size(); ++i)
        atol[i] = this->absolute_tolerance_[i];
      hip_detail::check(micm_hip_state_set_tolerances(device_.get(), this->relative_tolerance_, atol.data()),
                        "micm_hip_state_set_tolerances");
      hip_detail::check(micm_hip_state_set_lockstep_order(device_.get(), lockstep_order),
                        "micm_hip_state_set_lockstep_order");
    }

    /// Concentrations back to the host State.
    void SyncOutputsToHost()
    {
      const Index cells = this->number_of_grid_cells_;
      const Index cols = this->variables_.NumColumns();
      buffer_.resize(cells * cols);
      hip_detail::check(micm_hip_state_get_variables(device_.get(), buffer_.data()), "micm_hip_state_get_variables");
      for (Index c = 0; c < cells; ++c)
        for (Index i = 0; i < cols; ++i)
          this->variables_[c][i] = buffer_[c * cols + i];
    }
  };

  template<class RatesPolicy, class LinearSolverPolicy, class ConstraintSetPolicy>
  class HipRosenbrockSolver;

  /// SolverParametersPolicy of the HIP solver: the reference's parameters plus the form.
  struct HipRosenbrockSolverParameters : public RosenbrockSolverParameters
  {
    template<class RatesPolicy, class LinearSolverPolicy, class ConstraintSetPolicy>
    using SolverType = HipRosenbrockSolver<RatesPolicy, LinearSolverPolicy, ConstraintSetPolicy>;

    int32_t form_ = MICM_HIP_SOLVE_LOCKSTEP;

    HipRosenbrockSolverParameters(const RosenbrockSolverParameters& base, int32_t form = MICM_HIP_SOLVE_LOCKSTEP)
        : RosenbrockSolverParameters(base),
          form_(form)
    {
    }
  };

  /// SolverType: the whole Rosenbrock Solve (rosenbrock.inl:12-304) on the device.
  template<class RatesPolicy, class LinearSolverPolicy, class ConstraintSetPolicy>
  class HipRosenbrockSolver
  {
   public:
    using ParametersType = HipRosenbrockSolverParameters;

    LinearSolverPolicy linear_solver_;
    RatesPolicy rates_;
    ConstraintSetPolicy constraints_;

    HipRosenbrockSolver() = default;
    HipRosenbrockSolver(LinearSolverPolicy&& linear_solver, RatesPolicy&& rates, ConstraintSetPolicy&& constraints)
        : linear_solver_(std::move(linear_solver)),
          rates_(std::move(rates)),
          constraints_(std::move(constraints))
    {
      if (constraints_.Size() > 0)
        throw MicmException(MICM_ERROR_CATEGORY_SOLVER, MICM_SOLVER_ERROR_CODE_CUDA_CONSTRAINTS_UNSUPPORTED,
                            "HipRosenbrockSolver: constraints are not supported on the device");
    }
    HipRosenbrockSolver(HipRosenbrockSolver&&) noexcept = default;
    HipRosenbrockSolver& operator=(HipRosenbrockSolver&&) noexcept = default;

    template<class StatePolicy>
    SolverResult Solve(Real time_step, StatePolicy& state, const HipRosenbrockSolverParameters& parameters) const
    {
      using DenseMatrixPolicy = decltype(state.variables_);
      const micm_hip_mechanism* m = rates_.device_.get();
      state.SyncInputsToDevice(m, static_cast<int64_t>(DenseMatrixPolicy::GroupVectorSize()));
      const micm_hip_rosenbrock_parameters p = hip_detail::ToDevice(parameters);
      hip_detail::check(micm_hip_solve(m, state.Device(), time_step, &p, parameters.form_), "micm_hip_solve");
      state.SyncOutputsToHost();
      const Index cells = state.number_of_grid_cells_;
      std::vector<int32_t> status(cells);
      std::vector<int64_t> stats(cells * MICM_HIP_N_STATS);
      std::vector<double> final_time(cells);
      hip_detail::check(micm_hip_state_get_results(state.Device(), status.data(), stats.data(), final_time.data()),
                        "micm_hip_state_get_results");
      Index report = 0;
      for (Index c = 0; c < cells; ++c)
        if (status[c] != MICM_HIP_STATE_CONVERGED)
        {
          report = c;
          break;
        }
      SolverResult result;
      result.state_ = static_cast<SolverState>(status[report]);
      const int64_t* s = stats.data() + report * MICM_HIP_N_STATS;
      result.stats_.function_calls_ = static_cast<Index>(s[MICM_HIP_STAT_FUNCTION_CALLS]);
      result.stats_.jacobian_updates_ = static_cast<Index>(s[MICM_HIP_STAT_JACOBIAN_UPDATES]);
      result.stats_.number_of_steps_ = static_cast<Index>(s[MICM_HIP_STAT_STEPS]);
      result.stats_.accepted_ = static_cast<Index>(s[MICM_HIP_STAT_ACCEPTED]);
      result.stats_.rejected_ = static_cast<Index>(s[MICM_HIP_STAT_REJECTED]);
      result.stats_.decompositions_ = static_cast<Index>(s[MICM_HIP_STAT_DECOMPOSITIONS]);
      result.stats_.solves_ = static_cast<Index>(s[MICM_HIP_STAT_SOLVES]);
      result.stats_.final_time_ = final_time[report];
      return result;
    }
  };

  /// Builder of HIP solvers beside CpuSolverBuilder (solver_builder.hpp:224-236): vector-ordered
  /// host matrices with a group of L cells (64: one CDNA wavefront, the device layout).
  template<class SolverParametersPolicy, Index L = MICM_HIP_GROUP>
  using HipSolverBuilder = SolverBuilder<
      SolverParametersPolicy,
      VectorMatrix<Real, L>,
      SparseMatrix<Real, SparseMatrixVectorOrdering<L>>,
      HipProcessSet<VectorMatrix<Real, L>, SparseMatrix<Real, SparseMatrixVectorOrdering<L>>>,
      LuDecomposition<SparseMatrix<Real, SparseMatrixVectorOrdering<L>>>,
      LinearSolver<
          VectorMatrix<Real, L>,
          SparseMatrix<Real, SparseMatrixVectorOrdering<L>>,
          LuDecomposition<SparseMatrix<Real, SparseMatrixVectorOrdering<L>>>>,
      HipState<
          VectorMatrix<Real, L>,
          SparseMatrix<Real, SparseMatrixVectorOrdering<L>>,
          LuDecomposition<SparseMatrix<Real, SparseMatrixVectorOrdering<L>>>>>;
}  // namespace micm
