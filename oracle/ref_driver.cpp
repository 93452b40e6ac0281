// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Thin C-ABI driver around the reference's own header-only CPU solver
// (CpuSolverBuilder<RosenbrockSolverParameters>: Matrix<Real>, standard-ordering
// CSR, Doolittle LU, LinearSolver -- include/micm/solver/solver_builder.hpp:224-236).
// It is compiled against the reference headers where they lie (see
// oracle/Makefile, target `ref`) into oracle/_ref/libmicm_ref.so, and is used
// to pin the CPU restatement and to write golden fixtures.  Nothing on the
// product path links or loads it.
//
// Inputs use the flat description of include/micm_hip.h; every per-cell array is
// cell-major in the solver's state order ([cell][state index]) unless noted.

#include <micm/CPU.hpp>
#include <micm/process/chemical_reaction_builder.hpp>
#include <micm/solver/rosenbrock.hpp>
#include <micm/solver/solver_builder.hpp>

#include "../include/micm_hip.h"

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
  using Params = micm::RosenbrockSolverParameters;
  using Builder = micm::CpuSolverBuilder<Params>;
  using SolverT = decltype(std::declval<Builder&>().Build());
  using StateT = decltype(std::declval<SolverT&>().GetState(1));
  // the same solver over vector-ordered matrices (VectorMatrix<L>, SparseMatrixVectorOrdering<L>),
  // whose multi-cell State sums its error norm group by group
  constexpr std::size_t kRefVectorLength = 4;
  using BuilderV = micm::CpuSolverBuilder<Params, micm::VectorMatrix<double, kRefVectorLength>,
                                          micm::SparseMatrix<double, micm::SparseMatrixVectorOrdering<kRefVectorLength>>>;
  using SolverV = decltype(std::declval<BuilderV&>().Build());

  Params preset(int32_t p)
  {
    switch (p)
    {
      case 2: return Params::TwoStageRosenbrockParameters();
      case 3: return Params::ThreeStageRosenbrockParameters();
      case 4: return Params::FourStageRosenbrockParameters();
      case 41: return Params::FourStageDifferentialAlgebraicRosenbrockParameters();
      case 61: return Params::SixStageDifferentialAlgebraicRosenbrockParameters();
    }
    throw std::runtime_error("unknown tableau preset");
  }

  struct RefSolver
  {
    std::unique_ptr<SolverT> solver;
    std::unique_ptr<SolverV> solver_v;
    int64_t n_species = 0, n_reactions = 0, n_custom = 0;
    std::vector<std::string> species_names;          // declared order
    std::vector<std::string> custom_labels;          // in our custom-parameter order
    Params params = Params::ThreeStageRosenbrockParameters();
  };

  thread_local std::string g_error;

  micm::Process make_process(const micm_hip_mechanism_desc* d, int64_t r,
                             const std::vector<micm::Species>& species,
                             const std::vector<micm::Species>& parameterized,
                             const micm::Phase& phase, std::vector<std::string>& ud_labels,
                             std::vector<std::string>& surf_labels)
  {
    std::vector<micm::Species> reactants;
    for (int64_t i = d->reactant_offsets[r]; i < d->reactant_offsets[r + 1]; ++i)
    {
      int64_t id = d->reactants[i];
      reactants.push_back(id >= 0 ? species[id] : parameterized[-id - 1]);
    }
    std::vector<micm::StoichSpecies> products;
    for (int64_t i = d->product_offsets[r]; i < d->product_offsets[r + 1]; ++i)
    {
      int64_t id = d->products[i];
      products.push_back(micm::StoichSpecies(id >= 0 ? species[id] : parameterized[-id - 1], d->yields[i]));
    }
    const double* p = d->rate_params + r * MICM_HIP_RATE_PARAM_STRIDE;
    auto b = micm::ChemicalReactionBuilder().SetReactants(reactants).SetProducts(products).SetPhase(phase);
    switch (d->rate_types[r])
    {
      case MICM_HIP_RATE_ARRHENIUS:
        b.SetRateConstant(micm::ArrheniusRateConstantParameters{ .A_ = p[0], .B_ = p[1], .C_ = p[2], .D_ = p[3], .E_ = p[4] });
        break;
      case MICM_HIP_RATE_TROE:
        b.SetRateConstant(micm::TroeRateConstantParameters{ .k0_A_ = p[0], .k0_B_ = p[1], .k0_C_ = p[2], .kinf_A_ = p[3],
                                                            .kinf_B_ = p[4], .kinf_C_ = p[5], .Fc_ = p[6], .N_ = p[7] });
        break;
      case MICM_HIP_RATE_TERNARY:
        b.SetRateConstant(micm::TernaryChemicalActivationRateConstantParameters{
            .k0_A_ = p[0], .k0_B_ = p[1], .k0_C_ = p[2], .kinf_A_ = p[3], .kinf_B_ = p[4], .kinf_C_ = p[5], .Fc_ = p[6], .N_ = p[7] });
        break;
      case MICM_HIP_RATE_BRANCHED:
      {
        micm::BranchedRateConstantParameters bp{};
        bp.branch_ = p[0] == 0.0 ? micm::BranchedRateConstantParameters::Branch::Alkoxy
                                 : micm::BranchedRateConstantParameters::Branch::Nitrate;
        bp.X_ = p[1];
        bp.Y_ = p[2];
        bp.a0_ = p[3];
        bp.n_ = static_cast<micm::Index>(p[4]);
        b.SetRateConstant(bp);
        break;
      }
      case MICM_HIP_RATE_TUNNELING:
        b.SetRateConstant(micm::TunnelingRateConstantParameters{ .A_ = p[0], .B_ = p[1], .C_ = p[2] });
        break;
      case MICM_HIP_RATE_TAYLOR:
      {
        micm::TaylorSeriesRateConstantParameters tp{};
        tp.A_ = p[0];
        tp.B_ = p[1];
        tp.C_ = p[2];
        tp.D_ = p[3];
        tp.E_ = p[4];
        tp.n_coefficients_ = static_cast<micm::Index>(p[5]);
        for (int j = 0; j < 16; ++j)
          tp.coefficients_[j] = p[6 + j];
        b.SetRateConstant(tp);
        break;
      }
      case MICM_HIP_RATE_REVERSIBLE:
        b.SetRateConstant(micm::ReversibleRateConstantParameters{ .A_ = p[0], .C_ = p[1], .k_r_ = p[2] });
        break;
      case MICM_HIP_RATE_USER_DEFINED:
      {
        std::string label = "UD" + std::to_string(r);
        ud_labels[r] = label;
        b.SetRateConstant(micm::UserDefinedRateConstantParameters{ .label_ = label, .scaling_factor_ = p[0] });
        break;
      }
      case MICM_HIP_RATE_SURFACE:
      {
        std::string label = "SURF" + std::to_string(r);
        surf_labels[r] = label;
        micm::Species s("SURFSP" + std::to_string(r));
        s.SetProperty<double>(micm::property_keys::MOLECULAR_WEIGHT, p[1]);
        micm::PhaseSpecies ps(s, p[0]);
        b.SetRateConstant(micm::SurfaceRateConstantParameters{ .label_ = label, .phase_species_ = ps, .reaction_probability_ = p[2] });
        break;
      }
      default: throw std::runtime_error("unknown rate type");
    }
    return b.Build();
  }

  template<class Solver>
  auto make_state_of(RefSolver* s, Solver& solver, int64_t n_cells, const double* conditions, const double* custom,
                     double rtol, const double* atol)
  {
    auto state = solver.GetState(n_cells);
    for (int64_t c = 0; c < n_cells; ++c)
    {
      state.conditions_[c].temperature_ = conditions[c * 3 + 0];
      state.conditions_[c].pressure_ = conditions[c * 3 + 1];
      state.conditions_[c].air_density_ = conditions[c * 3 + 2];
      for (int64_t p = 0; p < s->n_custom; ++p)
        state.custom_rate_parameters_[c][state.custom_rate_parameter_map_.at(s->custom_labels[p])] = custom[c * s->n_custom + p];
    }
    if (rtol > 0)
      state.SetRelativeTolerance(rtol);
    if (atol)
    {
      std::vector<double> a(atol, atol + s->n_species);
      state.SetAbsoluteTolerances(a);
    }
    return state;
  }

  StateT make_state(RefSolver* s, int64_t n_cells, const double* conditions, const double* custom,
                    double rtol, const double* atol)
  {
    return make_state_of(s, *s->solver, n_cells, conditions, custom, rtol, atol);
  }
}  // namespace


namespace
{
  template<class Solver>
  int solve_with(RefSolver* s, Solver& solver, bool lockstep, int64_t n_cells, const double* conditions,
                 const double* custom, double* y, double dt, int64_t n_solves, double rtol, const double* atol,
                 int32_t* states, int64_t* stats, double* final_time)
  {
    const int64_t N = s->n_species;
    auto run = [&](int64_t c0, int64_t nc)
    {
      auto st = make_state_of(s, solver, nc, conditions + 3 * c0, custom + s->n_custom * c0, rtol, atol);
      for (int64_t c = 0; c < nc; ++c)
        for (int64_t i = 0; i < N; ++i)
          st.variables_[c][i] = y[(c0 + c) * N + i];
      std::vector<int64_t> acc(7, 0);
      micm::SolverResult res{};
      for (int64_t k = 0; k < n_solves; ++k)
      {
        solver.UpdateStateParameters(st);
        res = solver.Solve(dt, st);
        acc[0] += res.stats_.function_calls_;
        acc[1] += res.stats_.jacobian_updates_;
        acc[2] += res.stats_.number_of_steps_;
        acc[3] += res.stats_.accepted_;
        acc[4] += res.stats_.rejected_;
        acc[5] += res.stats_.decompositions_;
        acc[6] += res.stats_.solves_;
      }
      for (int64_t c = 0; c < nc; ++c)
      {
        for (int64_t i = 0; i < N; ++i)
          y[(c0 + c) * N + i] = st.variables_[c][i];
        if (states)
          states[c0 + c] = static_cast<int32_t>(res.state_);
        if (stats)
          for (int j = 0; j < 7; ++j)
            stats[(c0 + c) * 7 + j] = acc[j];
        if (final_time)
          final_time[c0 + c] = res.stats_.final_time_;
      }
    };
    try
    {
      if (lockstep)
        run(0, n_cells);
      else
        for (int64_t c = 0; c < n_cells; ++c)
          run(c, 1);
    }
    catch (const std::exception& e)
    {
      g_error = e.what();
      return -1;
    }
    return 0;
  }
}

extern "C"
{
  const char* ref_last_error()
  {
    return g_error.c_str();
  }

  void* ref_solver_create(const micm_hip_mechanism_desc* d, int32_t tableau)
  {
    try
    {
      auto s = std::make_unique<RefSolver>();
      s->n_species = d->n_species;
      s->n_reactions = d->n_reactions;
      std::vector<micm::Species> species;
      for (int64_t i = 0; i < d->n_species; ++i)
      {
        species.emplace_back("S" + std::to_string(i));
        if (d->absolute_tolerances)
          species.back().SetProperty<double>("absolute tolerance", d->absolute_tolerances[i]);
        s->species_names.push_back(species.back().name_);
      }
      std::vector<micm::Species> parameterized;
      for (int64_t i = 0; i < d->n_parameterized; ++i)
      {
        micm::Species p("P" + std::to_string(i));
        p.parameterize_ = { .c0_ = d->parameterized[4 * i + 0], .c_T_ = d->parameterized[4 * i + 1],
                            .c_P_ = d->parameterized[4 * i + 2], .c_rho_ = d->parameterized[4 * i + 3],
                            .has_value_ = static_cast<micm::Bool>(true) };
        parameterized.push_back(p);
      }
      std::vector<micm::PhaseSpecies> phase_species;
      for (auto& sp : species)
        phase_species.emplace_back(sp);
      for (auto& sp : parameterized)
        phase_species.emplace_back(sp);
      micm::Phase gas("gas", phase_species);
      std::vector<std::string> ud_labels(d->n_reactions), surf_labels(d->n_reactions);
      std::vector<micm::Process> processes;
      for (int64_t r = 0; r < d->n_reactions; ++r)
        processes.push_back(make_process(d, r, species, parameterized, gas, ud_labels, surf_labels));
      // custom parameters in sorted-reaction order: user-defined block first, then surface
      for (int type : { MICM_HIP_RATE_USER_DEFINED, MICM_HIP_RATE_SURFACE })
        for (int64_t r = 0; r < d->n_reactions; ++r)
          if (d->rate_types[r] == type)
          {
            if (type == MICM_HIP_RATE_USER_DEFINED)
              s->custom_labels.push_back(ud_labels[r]);
            else
            {
              s->custom_labels.push_back(surf_labels[r] + ".effective radius [m]");
              s->custom_labels.push_back(surf_labels[r] + ".particle number concentration [# m-3]");
            }
          }
      s->n_custom = (int64_t)s->custom_labels.size();
      s->params = preset(tableau);
      s->solver = std::make_unique<SolverT>(Builder(s->params)
                                                .SetSystem(micm::System(gas))
                                                .SetReactions(processes)
                                                .SetReorderState(d->reorder_state != 0)
                                                .SetIgnoreUnusedSpecies(d->ignore_unused_species != 0)
                                                .Build());
      s->solver_v = std::make_unique<SolverV>(BuilderV(s->params)
                                                  .SetSystem(micm::System(gas))
                                                  .SetReactions(processes)
                                                  .SetReorderState(d->reorder_state != 0)
                                                  .SetIgnoreUnusedSpecies(d->ignore_unused_species != 0)
                                                  .Build());
      return s.release();
    }
    catch (const std::exception& e)
    {
      g_error = e.what();
      return nullptr;
    }
  }

  // The reference's RosenbrockSolverParameters preset, every field, in the C-ABI's struct.
  int ref_tableau(int32_t preset_id, micm_hip_rosenbrock_parameters* out)
  {
    try
    {
      const Params p = preset(preset_id);
      *out = micm_hip_rosenbrock_parameters{};
      out->stages = static_cast<int64_t>(p.stages_);
      out->max_number_of_steps = static_cast<int64_t>(p.max_number_of_steps_);
      out->round_off = p.round_off_;
      out->factor_min = p.factor_min_;
      out->factor_max = p.factor_max_;
      out->rejection_factor_decrease = p.rejection_factor_decrease_;
      out->safety_factor = p.safety_factor_;
      out->h_min = p.h_min_;
      out->h_max = p.h_max_;
      out->h_start = p.h_start_;
      for (int i = 0; i < 6; ++i)
      {
        out->new_function_evaluation[i] = p.new_function_evaluation_[i] ? 1 : 0;
        out->m[i] = p.m_[i];
        out->e[i] = p.e_[i];
        out->alpha[i] = p.alpha_[i];
        out->gamma[i] = p.gamma_[i];
      }
      out->estimator_of_local_order = p.estimator_of_local_order_;
      for (int i = 0; i < 15; ++i)
      {
        out->a[i] = p.a_[i];
        out->c[i] = p.c_[i];
      }
      return 0;
    }
    catch (const std::exception& e)
    {
      g_error = e.what();
      return -1;
    }
  }

  void ref_solver_destroy(void* h)
  {
    delete static_cast<RefSolver*>(h);
  }

  // out: n_species, n_reactions, n_custom, nnz_jacobian, nnz_lower, nnz_upper
  int ref_sizes(void* h, int64_t* out)
  {
    auto* s = static_cast<RefSolver*>(h);
    StateT st = s->solver->GetState(1);
    out[0] = s->n_species;
    out[1] = s->n_reactions;
    out[2] = s->n_custom;
    out[3] = (int64_t)st.jacobian_.AsVector().size();
    out[4] = (int64_t)st.lower_matrix_.AsVector().size();
    out[5] = (int64_t)st.upper_matrix_.AsVector().size();
    return 0;
  }

  // state_to_declared[i] = declared species index held in state column i
  int ref_species_order(void* h, int64_t* state_to_declared)
  {
    auto* s = static_cast<RefSolver*>(h);
    StateT st = s->solver->GetState(1);
    for (int64_t i = 0; i < s->n_species; ++i)
      state_to_declared[st.variable_map_.at(s->species_names[i])] = i;
    return 0;
  }

  // (row, col) of each stored element of J, L and U in storage order
  int ref_patterns(void* h, int64_t* jr, int64_t* jc, int64_t* lr, int64_t* lc, int64_t* ur, int64_t* uc)
  {
    auto* s = static_cast<RefSolver*>(h);
    StateT st = s->solver->GetState(1);
    auto dump = [&](const auto& m, int64_t* rows, int64_t* cols)
    {
      const int64_t n = (int64_t)m.NumRows();
      for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j)
          if (!m.IsZero(i, j))
          {
            auto idx = m.VectorIndex(0, i, j);
            rows[idx] = i;
            cols[idx] = j;
          }
    };
    dump(st.jacobian_, jr, jc);
    dump(st.lower_matrix_, lr, lc);
    dump(st.upper_matrix_, ur, uc);
    return 0;
  }

  // rate constants [cell][sorted reaction] from conditions [cell][3] and custom [cell][P]
  int ref_rate_constants(void* h, int64_t n_cells, const double* conditions, const double* custom, double* k)
  {
    auto* s = static_cast<RefSolver*>(h);
    StateT st = make_state(s, n_cells, conditions, custom, 0, nullptr);
    s->solver->UpdateStateParameters(st);
    for (int64_t c = 0; c < n_cells; ++c)
      for (int64_t r = 0; r < s->n_reactions; ++r)
        k[c * s->n_reactions + r] = st.rate_constants_[c][r];
    return 0;
  }

  // forcing [cell][state] = AddForcingTerms on a zeroed matrix
  int ref_forcing(void* h, int64_t n_cells, const double* k, const double* y, double* f)
  {
    auto* s = static_cast<RefSolver*>(h);
    StateT st = s->solver->GetState(n_cells);
    auto Y = st.variables_;
    auto F = st.variables_;
    for (int64_t c = 0; c < n_cells; ++c)
    {
      for (int64_t r = 0; r < s->n_reactions; ++r)
        st.rate_constants_[c][r] = k[c * s->n_reactions + r];
      for (int64_t i = 0; i < s->n_species; ++i)
      {
        Y[c][i] = y[c * s->n_species + i];
        F[c][i] = 0.0;
      }
    }
    s->solver->solver_.rates_.AddForcingTerms(st, Y, F);
    for (int64_t c = 0; c < n_cells; ++c)
      for (int64_t i = 0; i < s->n_species; ++i)
        f[c * s->n_species + i] = F[c][i];
    return 0;
  }

  // AddForcingTerms into f [cell][species] as it holds (no Fill(0))
  int ref_add_forcing(void* h, int64_t n_cells, const double* k, const double* y, double* f)
  {
    auto* s = static_cast<RefSolver*>(h);
    StateT st = s->solver->GetState(n_cells);
    auto Y = st.variables_;
    auto F = st.variables_;
    for (int64_t c = 0; c < n_cells; ++c)
    {
      for (int64_t r = 0; r < s->n_reactions; ++r)
        st.rate_constants_[c][r] = k[c * s->n_reactions + r];
      for (int64_t i = 0; i < s->n_species; ++i)
      {
        Y[c][i] = y[c * s->n_species + i];
        F[c][i] = f[c * s->n_species + i];
      }
    }
    s->solver->solver_.rates_.AddForcingTerms(st, Y, F);
    for (int64_t c = 0; c < n_cells; ++c)
      for (int64_t i = 0; i < s->n_species; ++i)
        f[c * s->n_species + i] = F[c][i];
    return 0;
  }

  // SubtractJacobianTerms into jac [cell][nnz] as it holds (no Fill(0))
  int ref_subtract_jacobian(void* h, int64_t n_cells, const double* k, const double* y, double* jac)
  {
    auto* s = static_cast<RefSolver*>(h);
    StateT st = s->solver->GetState(n_cells);
    auto Y = st.variables_;
    for (int64_t c = 0; c < n_cells; ++c)
    {
      for (int64_t r = 0; r < s->n_reactions; ++r)
        st.rate_constants_[c][r] = k[c * s->n_reactions + r];
      for (int64_t i = 0; i < s->n_species; ++i)
        Y[c][i] = y[c * s->n_species + i];
    }
    auto& v = st.jacobian_.AsVector();
    std::memcpy(v.data(), jac, v.size() * sizeof(double));
    s->solver->solver_.rates_.SubtractJacobianTerms(st, Y, st.jacobian_);
    std::memcpy(jac, v.data(), v.size() * sizeof(double));
    return 0;
  }

  // -J [cell][nnz] = SubtractJacobianTerms on a zeroed matrix
  int ref_jacobian(void* h, int64_t n_cells, const double* k, const double* y, double* jac)
  {
    auto* s = static_cast<RefSolver*>(h);
    StateT st = s->solver->GetState(n_cells);
    auto Y = st.variables_;
    for (int64_t c = 0; c < n_cells; ++c)
    {
      for (int64_t r = 0; r < s->n_reactions; ++r)
        st.rate_constants_[c][r] = k[c * s->n_reactions + r];
      for (int64_t i = 0; i < s->n_species; ++i)
        Y[c][i] = y[c * s->n_species + i];
    }
    st.jacobian_.Fill(0);
    s->solver->solver_.rates_.SubtractJacobianTerms(st, Y, st.jacobian_);
    const auto& v = st.jacobian_.AsVector();
    std::memcpy(jac, v.data(), v.size() * sizeof(double));
    return 0;
  }

  // jac [cell][nnz] in: -J ; alpha applied on the diagonal (AlphaMinusJacobian), then Doolittle.
  int ref_lu(void* h, int64_t n_cells, const double* jac, double alpha, double* jac_out, double* lower, double* upper)
  {
    auto* s = static_cast<RefSolver*>(h);
    StateT st = s->solver->GetState(n_cells);
    auto& v = st.jacobian_.AsVector();
    std::memcpy(v.data(), jac, v.size() * sizeof(double));
    s->solver->solver_.template AlphaMinusJacobian<decltype(st.jacobian_)>(st, alpha);
    s->solver->solver_.linear_solver_.Factor(st.jacobian_, st.lower_matrix_, st.upper_matrix_);
    std::memcpy(jac_out, v.data(), v.size() * sizeof(double));
    std::memcpy(lower, st.lower_matrix_.AsVector().data(), st.lower_matrix_.AsVector().size() * sizeof(double));
    std::memcpy(upper, st.upper_matrix_.AsVector().data(), st.upper_matrix_.AsVector().size() * sizeof(double));
    return 0;
  }

  // x [cell][state] in/out
  int ref_lu_solve(void* h, int64_t n_cells, const double* lower, const double* upper, double* x)
  {
    auto* s = static_cast<RefSolver*>(h);
    StateT st = s->solver->GetState(n_cells);
    std::memcpy(st.lower_matrix_.AsVector().data(), lower, st.lower_matrix_.AsVector().size() * sizeof(double));
    std::memcpy(st.upper_matrix_.AsVector().data(), upper, st.upper_matrix_.AsVector().size() * sizeof(double));
    auto X = st.variables_;
    for (int64_t c = 0; c < n_cells; ++c)
      for (int64_t i = 0; i < s->n_species; ++i)
        X[c][i] = x[c * s->n_species + i];
    s->solver->solver_.linear_solver_.Solve(X, st.lower_matrix_, st.upper_matrix_);
    for (int64_t c = 0; c < n_cells; ++c)
      for (int64_t i = 0; i < s->n_species; ++i)
        x[c * s->n_species + i] = X[c][i];
    return 0;
  }

  // NormalizedError over one State holding n_cells (the reference's global norm)
  int ref_normalized_error(void* h, int64_t n_cells, const double* y, const double* ynew, const double* yerr,
                           double rtol, const double* atol, double* error)
  {
    auto* s = static_cast<RefSolver*>(h);
    StateT st = s->solver->GetState(n_cells);
    st.SetRelativeTolerance(rtol);
    std::vector<double> a(atol, atol + s->n_species);
    st.SetAbsoluteTolerances(a);
    auto Y = st.variables_, Yn = st.variables_, E = st.variables_;
    for (int64_t c = 0; c < n_cells; ++c)
      for (int64_t i = 0; i < s->n_species; ++i)
      {
        Y[c][i] = y[c * s->n_species + i];
        Yn[c][i] = ynew[c * s->n_species + i];
        E[c][i] = yerr[c * s->n_species + i];
      }
    typename decltype(Y)::template ScalarType<double> e{};
    s->solver->solver_.NormalizedError(Y, Yn, E, st, e);
    *error = static_cast<double>(e);
    return 0;
  }

  // n_solves consecutive (UpdateStateParameters, Solve(dt)) per cell.
  // lockstep = 1: one State holding every cell (the reference's multi-cell semantics);
  // lockstep = 4: the same over VectorMatrix<4> (the error norm summed group by group);
  // lockstep = 0: one 1-cell State per cell.  y is in/out.  states [cell] is the last
  // solve's SolverState, stats [cell][7] are summed over solves, final_time [cell] of the last solve.
  int ref_solve(void* h, int64_t n_cells, int32_t lockstep, const double* conditions, const double* custom,
                double* y, double dt, int64_t n_solves, double rtol, const double* atol, int32_t* states,
                int64_t* stats, double* final_time)
  {
    auto* s = static_cast<RefSolver*>(h);
    if (lockstep == static_cast<int32_t>(kRefVectorLength))
      return solve_with(s, *s->solver_v, true, n_cells, conditions, custom, y, dt, n_solves, rtol, atol, states,
                        stats, final_time);
    if (lockstep != 0 && lockstep != 1)
    {
      g_error = "lockstep: 0, 1 or " + std::to_string(kRefVectorLength);
      return -1;
    }
    return solve_with(s, *s->solver, lockstep != 0, n_cells, conditions, custom, y, dt, n_solves, rtol, atol, states,
                      stats, final_time);
  }
}
