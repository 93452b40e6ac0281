// TEST INFRASTRUCTURE: the README example of the reference (README.md:95-161) and a
// 1000-cell version of it run through micm::HipSolverBuilder (include/micm_hip_policies.hpp)
// and through the reference's own CpuSolverBuilder over the same vector-ordered matrices
// (VectorMatrix<double, 64>, SparseMatrixVectorOrdering<64>); the concentrations, states and
// statistics must agree bit for bit.  Also: the per-cell forms against one 1-cell CPU State
// per cell.  Built by oracle/Makefile (target hip_builder_test) against the reference
// headers; prints one line per check and exits non-zero on a mismatch.
#include "../../include/micm_hip_policies.hpp"

#include <micm/CPU.hpp>
#include <micm/process/chemical_reaction_builder.hpp>
#include <micm/process/rate_constant/arrhenius_rate_constant.hpp>
#include <micm/solver/rosenbrock.hpp>
#include <micm/solver/solver_builder.hpp>

#include <cstdio>
#include <cstring>
#include <iostream>
#include <random>
#include <vector>

using namespace micm;

namespace
{
  constexpr std::size_t L = 64;
  using CpuBuilderL = CpuSolverBuilder<RosenbrockSolverParameters, VectorMatrix<double, L>,
                                       SparseMatrix<double, SparseMatrixVectorOrdering<L>>>;

  struct System3
  {
    Species foo{ "Foo" }, bar{ "Bar" }, baz{ "Baz" };
    Phase gas{ "gas", std::vector<PhaseSpecies>{ foo, bar, baz } };
    std::vector<Process> reactions;
    System3()
    {
      reactions.push_back(ChemicalReactionBuilder()
                              .SetReactants({ foo })
                              .SetProducts({ StoichSpecies(bar, 0.8), StoichSpecies(baz, 0.2) })
                              .SetRateConstant(ArrheniusRateConstantParameters{ .A_ = 1.0e-3 })
                              .SetPhase(gas)
                              .Build());
      reactions.push_back(ChemicalReactionBuilder()
                              .SetReactants({ foo, bar })
                              .SetProducts({ StoichSpecies(baz, 1) })
                              .SetRateConstant(ArrheniusRateConstantParameters{ .A_ = 1.0e-5, .C_ = 110.0 })
                              .SetPhase(gas)
                              .Build());
    }
  };

  template<class StateT>
  void init_cells(StateT& state, const System3& s, std::size_t n, unsigned seed)
  {
    std::mt19937_64 rng(seed);
    for (std::size_t c = 0; c < n; ++c)
    {
      // the README's cell first, then temperatures 270-310 K and pressures 90-105 kPa
      const double u1 = (rng() >> 11) * 0x1.0p-53, u2 = (rng() >> 11) * 0x1.0p-53;
      state.conditions_[c].temperature_ = c == 0 ? 287.45 : 270.0 + 40.0 * u1;
      state.conditions_[c].pressure_ = c == 0 ? 101319.9 : 90.0e3 + 15.0e3 * u2;
      state.conditions_[c].CalculateIdealAirDensity();
      state.variables_[c][state.variable_map_.at("Foo")] = 20.0;
      state.variables_[c][state.variable_map_.at("Bar")] = 0.0;
      state.variables_[c][state.variable_map_.at("Baz")] = 0.0;
    }
  }

  bool same_bits(double a, double b)
  {
    return std::memcmp(&a, &b, sizeof a) == 0;
  }

  int failures = 0;

  void report(const char* what, bool ok)
  {
    std::printf("%s: %s\n", what, ok ? "ok" : "MISMATCH");
    if (!ok)
      ++failures;
  }

  // n cells in one State, 10 x 500 s: the HIP solver in `form` against the CPU solver
  // (one State for the lockstep form, one 1-cell State per cell for the per-cell forms)
  void compare(std::size_t n, int32_t form, const char* label)
  {
    System3 s;
    auto params = RosenbrockSolverParameters::ThreeStageRosenbrockParameters();
    auto hip = HipSolverBuilder<HipRosenbrockSolverParameters>(HipRosenbrockSolverParameters(params, form))
                   .SetSystem(System(s.gas))
                   .SetReactions(s.reactions)
                   .Build();
    auto cpu = CpuBuilderL(params).SetSystem(System(s.gas)).SetReactions(s.reactions).Build();
    auto hs = hip.GetState(n);
    init_cells(hs, s, n, 7);
    bool ok = true;
    std::vector<SolverResult> cpu_results;
    if (form == MICM_HIP_SOLVE_LOCKSTEP)
    {
      auto cs = cpu.GetState(n);
      init_cells(cs, s, n, 7);
      for (int i = 0; i < 10; ++i)
      {
        hip.UpdateStateParameters(hs);
        cpu.UpdateStateParameters(cs);
        auto rh = hip.Solve(500.0, hs);
        auto rc = cpu.Solve(500.0, cs);
        ok &= rh.state_ == rc.state_ && rh.stats_.number_of_steps_ == rc.stats_.number_of_steps_ &&
              rh.stats_.accepted_ == rc.stats_.accepted_ && rh.stats_.rejected_ == rc.stats_.rejected_ &&
              rh.stats_.function_calls_ == rc.stats_.function_calls_ && same_bits(rh.stats_.final_time_, rc.stats_.final_time_);
        for (std::size_t c = 0; c < n; ++c)
          for (std::size_t v = 0; v < 3; ++v)
            ok &= same_bits(hs.variables_[c][v], cs.variables_[c][v]);
        if (n == 1)
          std::printf("%5d, %10.2e, %10.2e, %10.2e\n", i * 500, hs.variables_[0][hs.variable_map_.at("Foo")],
                      hs.variables_[0][hs.variable_map_.at("Bar")], hs.variables_[0][hs.variable_map_.at("Baz")]);
      }
    }
    else
    {
      std::vector<decltype(cpu.GetState(1))> cells;
      for (std::size_t c = 0; c < n; ++c)
      {
        cells.push_back(cpu.GetState(1));
        auto tmp = cpu.GetState(n);
        init_cells(tmp, s, n, 7);
        cells.back().conditions_[0] = tmp.conditions_[c];
        for (std::size_t v = 0; v < 3; ++v)
          cells.back().variables_[0][v] = tmp.variables_[c][v];
      }
      for (int i = 0; i < 10; ++i)
      {
        hip.UpdateStateParameters(hs);
        auto rh = hip.Solve(500.0, hs);
        ok &= rh.state_ == SolverState::Converged;
        for (std::size_t c = 0; c < n; ++c)
        {
          cpu.UpdateStateParameters(cells[c]);
          auto rc = cpu.Solve(500.0, cells[c]);
          ok &= rc.state_ == SolverState::Converged;
          for (std::size_t v = 0; v < 3; ++v)
            ok &= same_bits(hs.variables_[c][v], cells[c].variables_[0][v]);
        }
      }
    }
    report(label, ok);
  }
}  // namespace

int main()
{
  try
  {
    compare(1, MICM_HIP_SOLVE_LOCKSTEP, "README example (1 cell, lockstep) bit for bit");
    compare(1000, MICM_HIP_SOLVE_LOCKSTEP, "1000 cells in one State (lockstep, VectorMatrix<64> order) bit for bit");
    compare(300, MICM_HIP_SOLVE_COOPERATIVE, "300 cells, cooperative form vs 1-cell CPU States bit for bit");
    compare(300, MICM_HIP_SOLVE_PERSISTENT, "300 cells, persistent form vs 1-cell CPU States bit for bit");
  }
  catch (const std::exception& e)
  {
    std::printf("exception: %s\n", e.what());
    return 2;
  }
  return failures == 0 ? 0 : 1;
}
