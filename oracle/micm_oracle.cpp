// micm_oracle.cpp -- TEST INFRASTRUCTURE ONLY: the CPU restatement of the
// reference's Rosenbrock path, used as the checker for the HIP kernels and as
// the "port" CPU baseline.  Nothing on the product path links or loads it.
//
// Written from the reference's algorithm description, one function per step:
//   reaction sort                      include/micm/solver/solver_builder.inl:274-329
//   species map / Markowitz reorder    solver_builder.inl:78-119, linear_solver.inl:29-124
//   ProcessSet tables                  include/micm/process/process_set.hpp:256-422
//   Jacobian pattern (+ diagonal)      process_set.hpp:425-457, include/micm/util/jacobian.hpp:15-32
//   Doolittle fill + numeric LU        include/micm/solver/lu_decomposition_doolittle.inl:69-175, 325-403
//   forward / back substitution        linear_solver.inl:235-294
//   rate constants                     include/micm/process/rate_constant/rate_constant_functions.hpp:51-136,
//                                      include/micm/process/reaction_rate_store.hpp:221-522
//   forcing / -Jacobian                process_set.hpp:524-670
//   Rosenbrock step control            include/micm/solver/rosenbrock.inl:12-402
//   post-solve clamp                   include/micm/solver/solver.hpp:275-297
//
// The symbolic parts use dense boolean N x N patterns (N <= a few hundred), a
// deliberately different formulation from the reference's sparse adjacency
// sets, so that a shared mistake is unlikely.  Every floating-point expression
// keeps the reference's operand order; the library is compiled with
// -ffp-contract=off so each +, -, *, / rounds as written.
//
// C-ABI: the same oracle_* functions as the ref_* functions of ref_driver.cpp;
// per-cell arrays are cell-major in the solver's state order.

#include "../include/micm_hip.h"

// ORACLE_PORTABLE_MATH: exp, log10 and pow of the rate constants and of the step
// controller come from micm_amd/csrc/pmath.h (the functions the HIP path uses)
// instead of the C library; everything else is unchanged.  This build isolates
// the device path's only deviation from the reference -- its math library.
#ifdef ORACLE_PORTABLE_MATH
#include "../micm_amd/csrc/pmath.h"
#define RC_EXP pmath::exp
#define RC_POW pmath::pow
#define RC_LOG10 pmath::log10
#else
#define RC_EXP std::exp
#define RC_POW std::pow
#define RC_LOG10 std::log10
#endif

#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
  thread_local std::string g_error;

  constexpr double kBoltzmann = 1.380649e-23;
  constexpr double kAvogadro = 6.02214076e23;
  constexpr double kGasConstant = kBoltzmann * kAvogadro;
  constexpr double kPi = 3.14159265358979323846;

  struct Tableau
  {
    int stages = 0;
    int64_t max_steps = 1000;
    double round_off = std::numeric_limits<double>::epsilon();
    double factor_min = 0.2, factor_max = 6, reject_decrease = 0.1, safety = 0.9;
    double h_min = 0, h_max = 0, h_start = 0;
    bool new_f[6] = {};
    double elo = 0;
    double a[15] = {}, c[15] = {}, m[6] = {}, e[6] = {}, gamma[6] = {};
  };

  // Tableau constants: Hairer & Wanner, Solving ODEs II (2/4/6-stage); Sandu et al.,
  // Atmos. Env. 31 (1997) (3-stage).  Values as rosenbrock_solver_parameters.hpp:152-433.
  Tableau make_tableau(int preset)
  {
    Tableau t;
    if (preset == 2)
    {
      double g = 1.0 + 1.0 / 1.41421356237309504880;
      t.stages = 2;
      t.a[0] = 1.0 / g;
      t.c[0] = -2.0 / g;
      t.new_f[0] = t.new_f[1] = true;
      t.m[0] = 3.0 / (2.0 * g);
      t.m[1] = 1.0 / (2.0 * g);
      t.e[0] = 1.0 / (2.0 * g);
      t.e[1] = 1.0 / (2.0 * g);
      t.elo = 2.0;
      t.gamma[0] = g;
      t.gamma[1] = -g;
    }
    else if (preset == 3)
    {
      t.stages = 3;
      t.a[0] = 1;
      t.a[1] = 1;
      t.c[0] = -0.10156171083877702091975600115545e+01;
      t.c[1] = 0.40759956452537699824805835358067e+01;
      t.c[2] = 0.92076794298330791242156818474003e+01;
      t.new_f[0] = t.new_f[1] = true;
      t.m[0] = 0.1e+01;
      t.m[1] = 0.61697947043828245592553615689730e+01;
      t.m[2] = -0.42772256543218573326238373806514;
      t.e[0] = 0.5;
      t.e[1] = -0.29079558716805469821718236208017e+01;
      t.e[2] = 0.22354069897811569627360909276199;
      t.elo = 3;
      t.gamma[0] = 0.43586652150845899941601945119356;
      t.gamma[1] = 0.24291996454816804366592249683314;
      t.gamma[2] = 0.21851380027664058511513169485832e+01;
    }
    else if (preset == 4)
    {
      t.stages = 4;
      t.a[0] = 0.2000000000000000E+01;
      t.a[1] = 0.1867943637803922E+01;
      t.a[2] = 0.2344449711399156;
      t.a[3] = t.a[1];
      t.a[4] = t.a[2];
      t.c[0] = -0.7137615036412310E+01;
      t.c[1] = 0.2580708087951457E+01;
      t.c[2] = 0.6515950076447975;
      t.c[3] = -0.2137148994382534E+01;
      t.c[4] = -0.3214669691237626;
      t.c[5] = -0.6949742501781779;
      t.new_f[0] = t.new_f[1] = t.new_f[2] = true;
      t.m[0] = 0.2255570073418735E+01;
      t.m[1] = 0.2870493262186792;
      t.m[2] = 0.4353179431840180;
      t.m[3] = 0.1093502252409163E+01;
      t.e[0] = -0.2815431932141155;
      t.e[1] = -0.7276199124938920E-01;
      t.e[2] = -0.1082196201495311;
      t.e[3] = -0.1093502252409163E+01;
      t.elo = 4.0;
      t.gamma[0] = 0.5728200000000000;
      t.gamma[1] = -0.1769193891319233E+01;
      t.gamma[2] = 0.7592633437920482;
      t.gamma[3] = -0.1049021087100450;
    }
    else if (preset == 41)
    {
      t.stages = 4;
      t.a[1] = 2.0;
      t.a[3] = 2.0;
      t.a[5] = 1.0;
      t.c[0] = 4.0;
      t.c[1] = 1.0;
      t.c[2] = -1.0;
      t.c[3] = 1.0;
      t.c[4] = -1.0;
      t.c[5] = -(8.0 / 3.0);
      t.new_f[0] = t.new_f[2] = t.new_f[3] = true;
      t.m[0] = 2.0;
      t.m[2] = 1.0;
      t.m[3] = 1.0;
      t.e[3] = 1.0;
      t.elo = 3.0;
      t.gamma[0] = 0.5;
      t.gamma[1] = 1.5;
    }
    else if (preset == 61)
    {
      t.stages = 6;
      t.gamma[0] = 0.25;
      t.gamma[1] = -0.1043;
      t.gamma[2] = 0.1035;
      t.gamma[3] = -0.3620000000000023E-01;
      const double a[15] = { 0.1544000000000000E+01, 0.9466785280815826, 0.2557011698983284, 0.3314825187068521E+01,
                             0.2896124015972201E+01, 0.9986419139977817, 0.1221224509226641E+01, 0.6019134481288629E+01,
                             0.1253708332932087E+02, -0.6878860361058950, 0.1221224509226641E+01, 0.6019134481288629E+01,
                             0.1253708332932087E+02, -0.6878860361058950, 1.0 };
      const double c[15] = { -0.5668800000000000E+01, -0.2430093356833875E+01, -0.2063599157091915,
                             -0.1073529058151375,     -0.9594562251023355E+01, -0.2047028614809616E+02,
                             0.7496443313967647E+01,  -0.1024680431464352E+02, -0.3399990352819905E+02,
                             0.1170890893206160E+02,  0.8083246795921522E+01,  -0.7981132988064893E+01,
                             -0.3152159432874371E+02, 0.1631930543123136E+02,  -0.6058818238834054E+01 };
      for (int i = 0; i < 15; ++i)
      {
        t.a[i] = a[i];
        t.c[i] = c[i];
      }
      t.m[0] = a[6];
      t.m[1] = a[7];
      t.m[2] = a[8];
      t.m[3] = a[9];
      t.m[4] = 1.0;
      t.m[5] = 1.0;
      t.e[5] = 1.0;
      for (int i = 0; i < 6; ++i)
        t.new_f[i] = true;
      t.elo = 4.0;
    }
    else
      throw std::runtime_error("unknown tableau preset");
    return t;
  }

  // ------------------------------------------------------------------ mechanism

  struct Reaction
  {
    int64_t declared = 0;
    int type = 0;
    std::array<double, MICM_HIP_RATE_PARAM_STRIDE> p{};
    std::vector<std::array<double, 4>> multipliers;  // parameterized reactants
    std::vector<int64_t> reactants;                  // state indices (parameterized skipped)
    std::vector<int64_t> products;
    std::vector<double> yields;
    int64_t custom = -1;  // first custom parameter
  };

  struct Mech
  {
    int64_t N = 0, R = 0, P = 0;
    std::vector<Reaction> rx;              // sorted order
    std::vector<int64_t> state_to_declared;
    std::vector<double> atol;              // state order
    Tableau tab;
    // patterns (state order), dense
    std::vector<char> A, Lp, Up;           // N*N
    std::vector<int64_t> a_pos, l_pos, u_pos;  // N*N -> storage index or -1
    int64_t nnzA = 0, nnzL = 0, nnzU = 0;
    std::vector<int64_t> diag;             // storage index of A[i][i]
    // Jacobian terms in the reference's order: (process, independent) pairs
    struct JacTerm
    {
      int64_t process, independent;
      std::vector<int64_t> dep_reactants;  // dependent reactant state ids
      std::vector<int64_t> flat;           // storage ids: deps..., independent, products...
    };
    std::vector<JacTerm> jac_terms;
    // Doolittle operation lists
    struct Elem
    {
      int64_t out;     // storage index in L or U
      int64_t a;       // storage index in A or -1
      int64_t div;     // U diagonal storage index for L elements, -1 for U
      std::vector<std::pair<int64_t, int64_t>> terms;  // (L index, U index), ascending j
    };
    std::vector<Elem> u_ops, l_ops;   // interleaved by row below
    std::vector<int64_t> row_u_count, row_l_count;
    std::vector<int64_t> l_diag;      // L[i][i] storage index
    // substitution lists
    std::vector<std::vector<std::pair<int64_t, int64_t>>> fwd, bwd;  // (matrix index, x index)
    std::vector<int64_t> u_diag;
  };

  std::vector<int64_t> markowitz(int64_t n, const std::vector<char>& pat)
  {
    // Dense restatement of DiagonalMarkowitzReorder (linear_solver.inl:29-124): pick the
    // live node with the smallest (row_deg-1)*(col_deg-1) in unsigned 64-bit arithmetic,
    // lowest index on ties; eliminating it adds an edge i->c for every live in-neighbour
    // i and out-neighbour c, and drops the pivot from its neighbours' degrees.
    std::vector<char> m(pat);
    std::vector<char> alive(n, 1);
    std::vector<int64_t> perm;
    for (int64_t step = 0; step < n; ++step)
    {
      int64_t pivot = -1;
      uint64_t best = 0;
      for (int64_t v = 0; v < n; ++v)
      {
        if (!alive[v])
          continue;
        uint64_t rd = 0, cd = 0;
        for (int64_t u = 0; u < n; ++u)
          if (alive[u] || u == v)
          {
            rd += m[v * n + u] && (alive[u] || u == v);
            cd += m[u * n + v] && (alive[u] || u == v);
          }
        uint64_t cost = (rd - 1) * (cd - 1);
        if (pivot < 0 || cost < best)
        {
          best = cost;
          pivot = v;
        }
      }
      perm.push_back(pivot);
      alive[pivot] = 0;
      for (int64_t i = 0; i < n; ++i)
      {
        if (!alive[i] || !m[i * n + pivot])
          continue;
        for (int64_t c = 0; c < n; ++c)
          if (alive[c] && m[pivot * n + c])
            m[i * n + c] = 1;
      }
    }
    return perm;
  }

  void build(Mech& M, const micm_hip_mechanism_desc* d, int tableau)
  {
    M.tab = make_tableau(tableau);
    M.N = d->n_species;
    M.R = d->n_reactions;
    const int64_t n = M.N;
    // ---- stable sort by type
    std::vector<int64_t> order(M.R);
    for (int64_t r = 0; r < M.R; ++r)
      order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return d->rate_types[x] < d->rate_types[y]; });
    // ---- pattern in declared indices (for the Markowitz order)
    std::vector<char> pat(n * n, 0);
    for (int64_t r = 0; r < M.R; ++r)
    {
      std::vector<int64_t> re, pr;
      for (int64_t i = d->reactant_offsets[r]; i < d->reactant_offsets[r + 1]; ++i)
        if (d->reactants[i] >= 0)
          re.push_back(d->reactants[i]);
      for (int64_t i = d->product_offsets[r]; i < d->product_offsets[r + 1]; ++i)
        if (d->products[i] >= 0)
          pr.push_back(d->products[i]);
      for (auto ind : re)
      {
        for (auto dep : re)
          pat[dep * n + ind] = 1;
        for (auto dep : pr)
          pat[dep * n + ind] = 1;
      }
    }
    if (!d->ignore_unused_species)
    {
      for (int64_t s = 0; s < n; ++s)
      {
        bool used = false;
        for (int64_t r = 0; r < M.R && !used; ++r)
        {
          for (int64_t i = d->reactant_offsets[r]; i < d->reactant_offsets[r + 1]; ++i)
            used |= d->reactants[i] == s;
          for (int64_t i = d->product_offsets[r]; i < d->product_offsets[r + 1]; ++i)
            used |= d->products[i] == s;
        }
        if (!used)
          throw std::runtime_error("Unused species in chemical system: S" + std::to_string(s));
      }
    }
    M.state_to_declared.resize(n);
    if (d->reorder_state)
      M.state_to_declared = markowitz(n, pat);
    else
      for (int64_t i = 0; i < n; ++i)
        M.state_to_declared[i] = i;
    std::vector<int64_t> declared_to_state(n);
    for (int64_t i = 0; i < n; ++i)
      declared_to_state[M.state_to_declared[i]] = i;
    M.atol.assign(n, 1e-3);
    if (d->absolute_tolerances)
      for (int64_t i = 0; i < n; ++i)
        M.atol[i] = d->absolute_tolerances[M.state_to_declared[i]];
    // ---- reactions in sorted order with state indices
    int64_t custom = 0;
    for (int64_t k = 0; k < M.R; ++k)
    {
      const int64_t r = order[k];
      Reaction rx;
      rx.declared = r;
      rx.type = d->rate_types[r];
      for (int j = 0; j < MICM_HIP_RATE_PARAM_STRIDE; ++j)
        rx.p[j] = d->rate_params[r * MICM_HIP_RATE_PARAM_STRIDE + j];
      for (int64_t i = d->reactant_offsets[r]; i < d->reactant_offsets[r + 1]; ++i)
      {
        int64_t id = d->reactants[i];
        if (id >= 0)
          rx.reactants.push_back(declared_to_state[id]);
        else
        {
          const double* f = d->parameterized + 4 * (-id - 1);
          rx.multipliers.push_back({ f[0], f[1], f[2], f[3] });
        }
      }
      for (int64_t i = d->product_offsets[r]; i < d->product_offsets[r + 1]; ++i)
        if (d->products[i] >= 0)
        {
          rx.products.push_back(declared_to_state[d->products[i]]);
          rx.yields.push_back(d->yields[i]);
        }
      if (rx.type == MICM_HIP_RATE_BRANCHED)
      {
        // derived fields of reaction_rate_store.hpp:286-293
        double nn = rx.p[4];
        double k0 = 2.0e-22 * kAvogadro * 1.0e-6 * std::exp(nn);
        double air_ref = 2.45e19 / kAvogadro * 1.0e6;
        double a = k0 * air_ref;
        double b = 0.43 * std::pow(293.0 / 298.0, -8.0);
        double A_val = a / (1.0 + a / b) * std::pow(0.41, 1.0 / (1.0 + std::pow(std::log10(a / b), 2.0)));
        rx.p[5] = k0;
        rx.p[6] = A_val * (1.0 - rx.p[3]) / rx.p[3];
      }
      if (rx.type == MICM_HIP_RATE_SURFACE)
        rx.p[3] = 8.0 * kGasConstant / (kPi * rx.p[1]);  // mean free speed factor
      if (rx.type == MICM_HIP_RATE_USER_DEFINED)
        rx.custom = custom++;
      else if (rx.type == MICM_HIP_RATE_SURFACE)
      {
        rx.custom = custom;
        custom += 2;
      }
      M.rx.push_back(rx);
    }
    M.P = custom;
    // ---- Jacobian pattern (state order) with the diagonal
    M.A.assign(n * n, 0);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < n; ++j)
        M.A[declared_to_state[i] * n + declared_to_state[j]] = pat[i * n + j];
    for (int64_t i = 0; i < n; ++i)
      M.A[i * n + i] = 1;
    // ---- Doolittle fill (dense restatement of lu_decomposition_doolittle.inl:69-175)
    M.Lp.assign(n * n, 0);
    M.Up.assign(n * n, 0);
    for (int64_t i = 0; i < n; ++i)
    {
      for (int64_t k = i; k < n; ++k)
      {
        bool nz = M.A[i * n + k] || k == i;
        for (int64_t j = 0; j < i && !nz; ++j)
          nz = M.Lp[i * n + j] && M.Up[j * n + k];
        M.Up[i * n + k] = nz;
      }
      M.Lp[i * n + i] = 1;
      for (int64_t k = i + 1; k < n; ++k)
      {
        bool nz = M.A[k * n + i];
        for (int64_t j = 0; j < i && !nz; ++j)
          nz = M.Lp[k * n + j] && M.Up[j * n + i];
        M.Lp[k * n + i] = nz;
      }
    }
    auto positions = [n](const std::vector<char>& p, std::vector<int64_t>& pos) -> int64_t
    {
      pos.assign(n * n, -1);
      int64_t c = 0;
      for (int64_t i = 0; i < n * n; ++i)
        if (p[i])
          pos[i] = c++;
      return c;
    };
    M.nnzA = positions(M.A, M.a_pos);
    M.nnzL = positions(M.Lp, M.l_pos);
    M.nnzU = positions(M.Up, M.u_pos);
    for (int64_t i = 0; i < n; ++i)
    {
      M.diag.push_back(M.a_pos[i * n + i]);
      M.l_diag.push_back(M.l_pos[i * n + i]);
      M.u_diag.push_back(M.u_pos[i * n + i]);
    }
    // ---- Jacobian terms: independent species ascending, then process order
    for (int64_t ind = 0; ind < n; ++ind)
      for (int64_t p = 0; p < M.R; ++p)
      {
        const auto& rx = M.rx[p];
        for (auto r : rx.reactants)
        {
          if (r != ind)
            continue;
          // one term per occurrence of `ind` among the reactants (2A gives two)
          Mech::JacTerm t;
          t.process = p;
          t.independent = ind;
          bool skipped = false;
          for (auto q : rx.reactants)
          {
            if (q == ind && !skipped)
            {
              skipped = true;
              continue;
            }
            t.dep_reactants.push_back(q);
          }
          for (auto q : t.dep_reactants)
            t.flat.push_back(M.a_pos[q * n + ind]);
          t.flat.push_back(M.a_pos[ind * n + ind]);
          for (auto q : rx.products)
            t.flat.push_back(M.a_pos[q * n + ind]);
          M.jac_terms.push_back(t);
        }
      }
    // ---- Doolittle operation lists (j ascending)
    for (int64_t i = 0; i < n; ++i)
    {
      int64_t cu = 0, cl = 0;
      for (int64_t k = i; k < n; ++k)
      {
        if (!M.Up[i * n + k])
          continue;
        Mech::Elem e;
        e.out = M.u_pos[i * n + k];
        e.a = M.a_pos[i * n + k];
        e.div = -1;
        for (int64_t j = 0; j < i; ++j)
          if (M.Lp[i * n + j] && M.Up[j * n + k])
            e.terms.push_back({ M.l_pos[i * n + j], M.u_pos[j * n + k] });
        M.u_ops.push_back(e);
        ++cu;
      }
      for (int64_t k = i + 1; k < n; ++k)
      {
        if (!M.Lp[k * n + i])
          continue;
        Mech::Elem e;
        e.out = M.l_pos[k * n + i];
        e.a = M.a_pos[k * n + i];
        e.div = M.u_pos[i * n + i];
        for (int64_t j = 0; j < i; ++j)
          if (M.Lp[k * n + j] && M.Up[j * n + i])
            e.terms.push_back({ M.l_pos[k * n + j], M.u_pos[j * n + i] });
        M.l_ops.push_back(e);
        ++cl;
      }
      M.row_u_count.push_back(cu);
      M.row_l_count.push_back(cl);
    }
    // ---- substitution lists
    M.fwd.assign(n, {});
    M.bwd.assign(n, {});
    for (int64_t i = 0; i < n; ++i)
    {
      for (int64_t j = 0; j < i; ++j)
        if (M.Lp[i * n + j])
          M.fwd[i].push_back({ M.l_pos[i * n + j], j });
      for (int64_t j = i + 1; j < n; ++j)
        if (M.Up[i * n + j])
          M.bwd[i].push_back({ M.u_pos[i * n + j], j });
    }
  }

  // ------------------------------------------------------------------ numerics (one cell)

  double rate_constant(const Reaction& rx, double T, double P, double rho, const double* custom)
  {
    const auto& p = rx.p;
    double k = 0;
    switch (rx.type)
    {
      case MICM_HIP_RATE_ARRHENIUS: k = p[0] * RC_EXP(p[2] / T) * RC_POW(T / p[3], p[1]) * (1.0 + p[4] * P); break;
      case MICM_HIP_RATE_TROE:
      case MICM_HIP_RATE_TERNARY:
      {
        double scale = rx.type == MICM_HIP_RATE_TROE ? rho : 1.0;
        double k0 = p[0] * RC_EXP(p[2] / T) * RC_POW(T / 300.0, p[1]);
        double kinf = p[3] * RC_EXP(p[5] / T) * RC_POW(T / 300.0, p[4]);
        double ratio = k0 * rho / kinf;
        // std::pow(x, 2.0) with a constant 2 compiles to x * x (rate_constant_functions.hpp:56)
        const double lg = RC_LOG10(ratio);
        k = k0 * scale / (1.0 + ratio) * RC_POW(p[6], p[7] / (p[7] + lg * lg));
        break;
      }
      case MICM_HIP_RATE_BRANCHED:
      {
        double a = p[5] * rho;
        double b = 0.43 * RC_POW(T / 298.0, -8.0);
        const double lg = RC_LOG10(a / b);  // std::pow(lg, 2.0) compiles to lg * lg
        double A_val = a / (1.0 + a / b) * RC_POW(0.41, 1.0 / (1.0 + lg * lg));
        double pre = p[1] * RC_EXP(-p[2] / T);
        k = p[0] == 0.0 ? pre * (p[6] / (p[6] + A_val)) : pre * (A_val / (A_val + p[6]));
        break;
      }
      case MICM_HIP_RATE_TUNNELING: k = p[0] * RC_EXP(-p[1] / T + p[2] / (T * T * T)); break;
      case MICM_HIP_RATE_TAYLOR:
      {
        double poly = 0.0, t_pow = 1.0;
        int64_t nc = static_cast<int64_t>(p[5]);
        for (int64_t j = 0; j < nc; ++j, t_pow *= T)
          poly += p[6 + j] * t_pow;
        k = poly * p[0] * RC_EXP(p[2] / T) * RC_POW(T / p[3], p[1]) * (1.0 + p[4] * P);
        break;
      }
      case MICM_HIP_RATE_REVERSIBLE: k = p[0] * RC_EXP(p[1] / T) * p[2]; break;
      case MICM_HIP_RATE_USER_DEFINED: k = custom[rx.custom] * p[0]; break;
      case MICM_HIP_RATE_SURFACE:
      {
        double radius = custom[rx.custom], num = custom[rx.custom + 1];
        double mfs = std::sqrt(p[3] * T);
        k = 4.0 * num * kPi * radius * radius / (radius / p[0] + 4.0 / (mfs * p[2]));
        break;
      }
    }
    if (!rx.multipliers.empty())
    {
      double v = 1.0;
      for (const auto& f : rx.multipliers)
        v *= f[0] + f[1] * T + f[2] * P + f[3] * rho;
      k *= v;
    }
    return k;
  }

  void add_forcing(const Mech& M, const double* k, const double* y, double* f)
  {
    for (int64_t r = 0; r < M.R; ++r)
    {
      const auto& rx = M.rx[r];
      double rate = k[r];
      for (auto s : rx.reactants)
        rate *= y[s];
      for (auto s : rx.reactants)
        f[s] -= rate;
      for (size_t j = 0; j < rx.products.size(); ++j)
        f[rx.products[j]] += rx.yields[j] * rate;
    }
  }

  void subtract_jacobian(const Mech& M, const double* k, const double* y, double* jac)
  {
    for (const auto& t : M.jac_terms)
    {
      const auto& rx = M.rx[t.process];
      double d = k[t.process];
      for (auto q : t.dep_reactants)
        d *= y[q];
      size_t f = 0;
      for (size_t j = 0; j < t.dep_reactants.size(); ++j)
        jac[t.flat[f++]] += d;
      jac[t.flat[f++]] += d;
      for (size_t j = 0; j < rx.products.size(); ++j)
        jac[t.flat[f++]] -= rx.yields[j] * d;
    }
  }

  void lu_decompose(const Mech& M, const double* A, double* L, double* U)
  {
    size_t iu = 0, il = 0;
    for (int64_t i = 0; i < M.N; ++i)
    {
      for (int64_t c = 0; c < M.row_u_count[i]; ++c)
      {
        const auto& e = M.u_ops[iu++];
        double u = e.a >= 0 ? A[e.a] : 0.0;
        for (const auto& t : e.terms)
          u -= L[t.first] * U[t.second];
        U[e.out] = u;
      }
      L[M.l_diag[i]] = 1.0;
      for (int64_t c = 0; c < M.row_l_count[i]; ++c)
      {
        const auto& e = M.l_ops[il++];
        double l = e.a >= 0 ? A[e.a] : 0.0;
        for (const auto& t : e.terms)
          l -= L[t.first] * U[t.second];
        l /= U[e.div];
        L[e.out] = l;
      }
    }
  }

  void lu_solve(const Mech& M, const double* L, const double* U, double* x)
  {
    for (int64_t i = 0; i < M.N; ++i)
    {
      for (const auto& t : M.fwd[i])
        x[i] -= L[t.first] * x[t.second];
      x[i] /= L[M.l_diag[i]];
    }
    for (int64_t i = M.N - 1; i >= 0; --i)
    {
      for (const auto& t : M.bwd[i])
        x[i] -= U[t.first] * x[t.second];
      x[i] /= U[M.u_diag[i]];
    }
  }

  struct Stats
  {
    int64_t v[7] = {};
    double final_time = 0;
    int state = MICM_HIP_STATE_NOT_YET_CALLED;
  };

  // One Rosenbrock Solve over `nc` cells sharing one step size (rosenbrock.inl:12-304,
  // separate-LU path), followed by the post-solve clamp.  nc == 1 gives the per-cell form.
  //
  // With `tape` (nc == 1) the Solve appends its records for the discrete adjoint: one per
  // accepted step (H, then y at the step's start) and one for the clamp (kClampTag, then y
  // before the clamp); each record is N + 1 doubles.
  constexpr double kClampTag = -1.0;

  // `group`: vector length of the State's dense matrix policy, which fixes the order of the
  // error norm's sum -- 1 for Matrix<Real> (cells outer, species inner), L for VectorMatrix<L>
  // (groups of L cells, species, then the group's cells; vector_matrix.hpp:860-867).
  void solve_state(const Mech& M, int64_t nc, const double* k, double* Y, double dt, double rtol,
                   const double* atol, Stats& st, std::vector<double>* tape = nullptr, int64_t group = 1)
  {
    const Tableau& p = M.tab;
    const int64_t N = M.N, nJ = M.nnzA;
    std::vector<double> Ynew(nc * N), F0(nc * N), Yerr(nc * N), J(nc * nJ), L(nc * M.nnzL), U(nc * M.nnzU);
    std::vector<std::vector<double>> K(p.stages, std::vector<double>(nc * N));
    auto forcing = [&](const double* y, double* f)
    {
      for (int64_t c = 0; c < nc * N; ++c)
        f[c] = 0.0;
      for (int64_t c = 0; c < nc; ++c)
        add_forcing(M, k + c * M.R, y + c * N, f + c * N);
    };
    auto jacobian = [&]()
    {
      std::fill(J.begin(), J.end(), 0.0);
      for (int64_t c = 0; c < nc; ++c)
        subtract_jacobian(M, k + c * M.R, Y + c * N, J.data() + c * nJ);
    };
    auto axpy = [&](double* y, double a, const double* x)
    {
      for (int64_t c = 0; c < nc * N; ++c)
        y[c] += a * x[c];
    };

    const double h_min = p.h_min == 0.0 ? 1.0e-15 * dt : p.h_min;
    const double h_max = p.h_max == 0.0 ? dt : std::min(dt, p.h_max);
    const double h_start = p.h_start == 0.0 ? 1.0e-6 * dt : std::min(h_max, p.h_start);
    double H = std::min(std::max(h_min, std::abs(h_start)), std::abs(h_max));
    double t = 0.0;
    bool reject_last = false, reject_more = false;
    int state = MICM_HIP_STATE_RUNNING;
    int64_t* s = st.v;
    std::fill(s, s + 7, 0);
    std::vector<double>* Ycur = nullptr;  // Y lives in the caller's buffer; Ynew swaps by copy
    (void)Ycur;
    while ((t - dt + p.round_off) <= 0 && state == MICM_HIP_STATE_RUNNING)
    {
      if (s[MICM_HIP_STAT_STEPS] > p.max_steps)
      {
        state = MICM_HIP_STATE_MAX_STEPS;
        break;
      }
      if (((t + 0.1 * H) == t) || (H <= p.round_off))
      {
        state = MICM_HIP_STATE_STEP_SIZE_TOO_SMALL;
        break;
      }
      H = std::min(H, std::abs(dt - t));
      forcing(Y, F0.data());
      s[MICM_HIP_STAT_FUNCTION_CALLS] += 1;
      jacobian();
      s[MICM_HIP_STAT_JACOBIAN_UPDATES] += 1;
      bool accepted = false;
      double last_alpha = 0.0;
      while (!accepted)
      {
        double alpha = 1.0 / (H * p.gamma[0]);
        const double cumulative = alpha;
        alpha -= last_alpha;
        last_alpha = cumulative;
        for (int64_t c = 0; c < nc; ++c)
          for (int64_t i = 0; i < N; ++i)
            J[c * nJ + M.diag[i]] += alpha * 1.0;
        for (int64_t c = 0; c < nc; ++c)
          lu_decompose(M, J.data() + c * nJ, L.data() + c * M.nnzL, U.data() + c * M.nnzU);
        s[MICM_HIP_STAT_DECOMPOSITIONS] += 1;
        for (int stage = 0; stage < p.stages; ++stage)
        {
          const int sc = stage * (stage - 1) / 2;
          if (stage == 0)
            K[0] = F0;
          else if (p.new_f[stage])
          {
            std::memcpy(Ynew.data(), Y, sizeof(double) * nc * N);
            for (int j = 0; j < stage; ++j)
              axpy(Ynew.data(), p.a[sc + j], K[j].data());
            forcing(Ynew.data(), K[stage].data());
            s[MICM_HIP_STAT_FUNCTION_CALLS] += 1;
          }
          if (stage + 1 < p.stages && !p.new_f[stage + 1])
            K[stage + 1] = K[stage];
          for (int j = 0; j < stage; ++j)
            axpy(K[stage].data(), p.c[sc + j] / H, K[j].data());
          for (int64_t c = 0; c < nc; ++c)
            lu_solve(M, L.data() + c * M.nnzL, U.data() + c * M.nnzU, K[stage].data() + c * N);
          s[MICM_HIP_STAT_SOLVES] += 1;
        }
        std::memcpy(Ynew.data(), Y, sizeof(double) * nc * N);
        for (int stage = 0; stage < p.stages; ++stage)
          axpy(Ynew.data(), p.m[stage], K[stage].data());
        std::fill(Yerr.begin(), Yerr.end(), 0.0);
        for (int stage = 0; stage < p.stages; ++stage)
          axpy(Yerr.data(), p.e[stage], K[stage].data());
        // NormalizedError in the State's summation order
        double error = 0;
        for (int64_t g0 = 0; g0 < nc; g0 += group)
        {
          const int64_t rows = std::min(group, nc - g0);
          for (int64_t i = 0; i < N; ++i)
            for (int64_t c = g0; c < g0 + rows; ++c)
            {
              double y = Y[c * N + i], yn = Ynew[c * N + i];
              double ymax = (std::abs(y) > std::abs(yn) ? std::abs(y) : std::abs(yn));
              double q = Yerr[c * N + i] / (atol[i] + rtol * ymax);
              error += q * q;
            }
        }
        const double Ntot = static_cast<double>(std::max<int64_t>(1, nc * N));
        error = std::max(std::sqrt(error / Ntot), 1.0e-10);
        double fac = std::min(p.factor_max, std::max(p.factor_min, p.safety / RC_POW(error, 1 / p.elo)));
        double Hnew = H * fac;
        s[MICM_HIP_STAT_STEPS] += 1;
        if (std::isnan(error))
        {
          state = MICM_HIP_STATE_NAN_DETECTED;
          break;
        }
        if (std::isinf(error))
        {
          state = MICM_HIP_STATE_INF_DETECTED;
          break;
        }
        if ((error < 1) || (H < h_min))
        {
          s[MICM_HIP_STAT_ACCEPTED] += 1;
          if (tape)
          {
            tape->push_back(H);
            tape->insert(tape->end(), Y, Y + N);
          }
          t = t + H;
          std::memcpy(Y, Ynew.data(), sizeof(double) * nc * N);
          Hnew = std::max(h_min, std::min(Hnew, h_max));
          if (reject_last)
            Hnew = std::min(Hnew, H);
          reject_last = false;
          reject_more = false;
          H = Hnew;
          accepted = true;
        }
        else
        {
          if (reject_more)
            Hnew = H * p.reject_decrease;
          reject_more = reject_last;
          reject_last = true;
          H = Hnew;
          if (s[MICM_HIP_STAT_ACCEPTED] >= 1)
            s[MICM_HIP_STAT_REJECTED] += 1;
        }
      }
    }
    if (state == MICM_HIP_STATE_RUNNING)
      state = MICM_HIP_STATE_CONVERGED;
    st.final_time = t;
    st.state = state;
    if (tape)
    {
      tape->push_back(kClampTag);
      tape->insert(tape->end(), Y, Y + N);
    }
    for (int64_t c = 0; c < nc * N; ++c)
      Y[c] = std::max(Y[c], 0.0);  // PostSolveClamp (solver.hpp:295)
  }

  // ------------------------------------------------------------------ discrete adjoint
  //
  // The adjoint differentiates the map the recorded Solve applied: each accepted step with
  // its step size H frozen (the reference has no adjoint; the step is rosenbrock.inl:140-207
  // with A = alpha I - J(y), alpha = 1 / (H gamma_0), and the clamp of solver.hpp:275-297
  // after each Solve).  Rejected attempts only chose H, so they drop out.
  //
  // One accepted step, S stages, f-evaluations g_s = f(Y_s) at Y_0 = y and
  // Y_s = y + sum_{j<s} a_sj K_j where new_function_evaluation[s], else g_s = g_{s-1}:
  //     r_s = g_s + sum_{j<s} (c_sj / H) K_j,   K_s = A^-1 r_s,   ynew = y + sum_s m_s K_s.
  // Reverse, with lam = dL/dynew:  Kbar_s = m_s lam;  for s = S-1 .. 0:
  //     rbar = A^-T Kbar_s;  Jbar += rbar K_s^T (A = alpha I - J);  gbar += rbar;
  //     Kbar_j += (c_sj / H) rbar;  at an f-evaluation: ybar_s = J(Y_s)^T gbar,
  //     lam += ybar_s, Kbar_j += a_sj ybar_s, gbar = 0.
  // Jbar reaches y and k through the second derivative of the mass-action rates.
  // Rate parameters enter as per-cell log k (d/dlog k = k d/dk), which for every rate type
  // is the derivative with respect to the log of a common factor on its A-prefactors.

  // Planted mistakes for the tests that show the adjoint's checks are sharp (0: none;
  // 1: drop the Jacobian's dependence on y and k; 2: solve with A instead of A^T;
  // 3: flip the sign of the forcing's k-derivative).
  int g_adjoint_fault = 0;

  struct StepData
  {
    std::vector<double> J, L, U, K, Ys, f, v;
  };

  void step_data_init(const Mech& M, StepData& w)
  {
    const int64_t S = M.tab.stages;
    w.J.assign(M.nnzA, 0.0);
    w.L.assign(M.nnzL, 0.0);
    w.U.assign(M.nnzU, 0.0);
    w.K.assign(S * M.N, 0.0);
    w.Ys.assign(S * M.N, 0.0);
    w.f.assign(M.N, 0.0);
    w.v.assign(std::max(M.N, M.nnzA), 0.0);
  }

  // One accepted step with a frozen H: the stage recurrences of rosenbrock.inl:140-207 on
  // A = alpha I - J(y); keeps the factors, the solved K_s and the forcing inputs Y_s.
  void frozen_step(const Mech& M, const double* k, const double* y, double H, StepData& w, double* ynew)
  {
    const Tableau& p = M.tab;
    const int64_t N = M.N;
    std::fill(w.J.begin(), w.J.end(), 0.0);
    subtract_jacobian(M, k, y, w.J.data());
    const double alpha = 1.0 / (H * p.gamma[0]);
    for (int64_t i = 0; i < N; ++i)
      w.J[M.diag[i]] += alpha * 1.0;
    lu_decompose(M, w.J.data(), w.L.data(), w.U.data());
    for (int s = 0; s < p.stages; ++s)
    {
      const int sc = s * (s - 1) / 2;
      double* Ks = w.K.data() + s * N;
      double* Ysv = w.Ys.data() + s * N;
      if (s == 0 || p.new_f[s])
      {
        for (int64_t i = 0; i < N; ++i)
        {
          double v = y[i];
          for (int j = 0; j < s; ++j)
            v += p.a[sc + j] * w.K[j * N + i];
          Ysv[i] = s == 0 ? y[i] : v;
        }
        std::fill(w.f.begin(), w.f.end(), 0.0);
        add_forcing(M, k, Ysv, w.f.data());
      }
      else
      {
        std::memcpy(Ysv, Ysv - N, sizeof(double) * N);
      }
      for (int64_t i = 0; i < N; ++i)
      {
        double v = w.f[i];
        for (int j = 0; j < s; ++j)
          v += (p.c[sc + j] / H) * w.K[j * N + i];
        Ks[i] = v;
      }
      lu_solve(M, w.L.data(), w.U.data(), Ks);
    }
    for (int64_t i = 0; i < N; ++i)
    {
      double v = y[i];
      for (int s = 0; s < p.stages; ++s)
        v += p.m[s] * w.K[s * N + i];
      ynew[i] = v;
    }
  }

  // x <- A^-T x with A = L U: U^T z = x (rows ascending), then L^T x = z (rows descending)
  void lu_solve_transposed(const Mech& M, const double* L, const double* U, double* x)
  {
    const int64_t n = M.N;
    for (int64_t i = 0; i < n; ++i)
    {
      for (int64_t j = 0; j < i; ++j)
        if (M.u_pos[j * n + i] >= 0)
          x[i] -= U[M.u_pos[j * n + i]] * x[j];
      x[i] /= U[M.u_diag[i]];
    }
    for (int64_t i = n - 1; i >= 0; --i)
    {
      for (int64_t j = i + 1; j < n; ++j)
        if (M.l_pos[j * n + i] >= 0)
          x[i] -= L[M.l_pos[j * n + i]] * x[j];
      x[i] /= L[M.l_diag[i]];
    }
  }

  // The Jacobian term of `t` adds d_t = k_r prod_{deps} y to -J[dep rows and the
  // independent row q, q] and -yield d_t to -J[product rows, q] (process_set.hpp:594-670).
  // Its share of u^T J v is therefore d_t v_q (sum yield u_prod - sum u_dep - u_q).
  double term_weight(const Mech& M, const Mech::JacTerm& t, const double* u, const double* v)
  {
    const auto& rx = M.rx[t.process];
    double w = 0.0;
    for (size_t j = 0; j < rx.products.size(); ++j)
      w += rx.yields[j] * u[rx.products[j]];
    for (auto q : t.dep_reactants)
      w -= u[q];
    w -= u[t.independent];
    return w * v[t.independent];
  }

  // ybar += d/dy (u^T J(y) v), th += k d/dk (u^T J(y) v)
  void jacobian_second_order(const Mech& M, const double* k, const double* y, const double* u, const double* v,
                             double* ybar, double* th)
  {
    for (const auto& t : M.jac_terms)
    {
      const double s = term_weight(M, t, u, v);
      const size_t nd = t.dep_reactants.size();
      double d = k[t.process];
      for (auto q : t.dep_reactants)
        d *= y[q];
      th[t.process] += s * d;
      for (size_t a = 0; a < nd; ++a)
      {
        double part = k[t.process];
        for (size_t b = 0; b < nd; ++b)
          if (b != a)
            part *= y[t.dep_reactants[b]];
        ybar[t.dep_reactants[a]] += s * part;
      }
    }
  }

  // out = J(y)^T u from the assembled -J (a different route from the rates' derivatives);
  // th += k d/dk (u^T f(y)) = rate_r (sum yield u_prod - sum u_reactant)
  void forcing_vjp(const Mech& M, const double* k, const double* y, const double* u, StepData& w, double* out,
                   double* th)
  {
    const int64_t n = M.N;
    std::vector<double> mj(M.nnzA, 0.0);
    subtract_jacobian(M, k, y, mj.data());
    for (int64_t q = 0; q < n; ++q)
    {
      double acc = 0.0;
      for (int64_t i = 0; i < n; ++i)
        if (M.a_pos[i * n + q] >= 0)
          acc -= mj[M.a_pos[i * n + q]] * u[i];
      out[q] = acc;
    }
    (void)w;
    for (int64_t r = 0; r < M.R; ++r)
    {
      const auto& rx = M.rx[r];
      double rate = k[r];
      for (auto s : rx.reactants)
        rate *= y[s];
      double wr = 0.0;
      for (size_t j = 0; j < rx.products.size(); ++j)
        wr += rx.yields[j] * u[rx.products[j]];
      for (auto s : rx.reactants)
        wr -= u[s];
      th[r] += (g_adjoint_fault == 3 ? -rate : rate) * wr;
    }
  }

  // Adjoint of one accepted step: lam in dL/dynew, out dL/dy; th += k dL/dk.
  void step_adjoint(const Mech& M, const double* k, const double* y, double H, StepData& w, double* lam, double* th)
  {
    const Tableau& p = M.tab;
    const int64_t N = M.N;
    const int S = p.stages;
    std::vector<double> ynew(N), Kb(S * N), gb(N, 0.0), rb(N), yb(N);
    frozen_step(M, k, y, H, w, ynew.data());
    for (int s = 0; s < S; ++s)
      for (int64_t i = 0; i < N; ++i)
        Kb[s * N + i] = p.m[s] * lam[i];
    for (int s = S - 1; s >= 0; --s)
    {
      const int sc = s * (s - 1) / 2;
      std::memcpy(rb.data(), Kb.data() + s * N, sizeof(double) * N);
      if (g_adjoint_fault == 2)
        lu_solve(M, w.L.data(), w.U.data(), rb.data());
      else
        lu_solve_transposed(M, w.L.data(), w.U.data(), rb.data());
      if (g_adjoint_fault != 1)
        jacobian_second_order(M, k, y, rb.data(), w.K.data() + s * N, lam, th);
      for (int64_t i = 0; i < N; ++i)
        gb[i] += rb[i];
      for (int j = 0; j < s; ++j)
        for (int64_t i = 0; i < N; ++i)
          Kb[j * N + i] += (p.c[sc + j] / H) * rb[i];
      if (s == 0 || p.new_f[s])
      {
        forcing_vjp(M, k, w.Ys.data() + s * N, gb.data(), w, yb.data(), th);
        for (int64_t i = 0; i < N; ++i)
          lam[i] += yb[i];
        for (int j = 0; j < s; ++j)
          for (int64_t i = 0; i < N; ++i)
            Kb[j * N + i] += p.a[sc + j] * yb[i];
        std::fill(gb.begin(), gb.end(), 0.0);
      }
    }
  }

  // Tangent of one accepted step along nd directions: dy [nd][N] in/out, dth [nd][R]
  // (log k).  A different formulation from the adjoint: derivatives of the assembled
  // -J and of the forcing, solved with A (not its transpose).
  void step_tangent(const Mech& M, const double* k, double* y, double H, StepData& w, int64_t nd, double* dy,
                    const double* dth)
  {
    const Tableau& p = M.tab;
    const int64_t N = M.N, n = M.N;
    const int S = p.stages;
    std::vector<double> ynew(N), mj(M.nnzA), dK(S * N), dg(N), dr(N), dys(N);
    frozen_step(M, k, y, H, w, ynew.data());
    for (int64_t d = 0; d < nd; ++d)
    {
      double* dyd = dy + d * N;
      const double* dt = dth + d * M.R;
      for (int s = 0; s < S; ++s)
      {
        const int sc = s * (s - 1) / 2;
        const double* ys = w.Ys.data() + s * N;
        if (s == 0 || p.new_f[s])
        {
          for (int64_t i = 0; i < N; ++i)
          {
            double v = dyd[i];
            for (int j = 0; j < s; ++j)
              v += p.a[sc + j] * dK[j * N + i];
            dys[i] = v;
          }
          // J(Y_s) dys from the assembled -J, plus the rates' log-k derivatives
          std::fill(mj.begin(), mj.end(), 0.0);
          subtract_jacobian(M, k, ys, mj.data());
          for (int64_t i = 0; i < n; ++i)
          {
            double acc = 0.0;
            for (int64_t q = 0; q < n; ++q)
              if (M.a_pos[i * n + q] >= 0)
                acc -= mj[M.a_pos[i * n + q]] * dys[q];
            dg[i] = acc;
          }
          for (int64_t r = 0; r < M.R; ++r)
          {
            const auto& rx = M.rx[r];
            double rate = k[r] * dt[r];
            for (auto q : rx.reactants)
              rate *= ys[q];
            for (auto q : rx.reactants)
              dg[q] -= rate;
            for (size_t j = 0; j < rx.products.size(); ++j)
              dg[rx.products[j]] += rx.yields[j] * rate;
          }
        }
        for (int64_t i = 0; i < N; ++i)
        {
          double v = dg[i];
          for (int j = 0; j < s; ++j)
            v += (p.c[sc + j] / H) * dK[j * N + i];
          dr[i] = v;
        }
        // -dA K_s = dJ K_s: each Jacobian term's d_t changes by
        // dd_t = d_t dth_r + k_r sum_a (prod_{deps but a} y) dy_a
        const double* Ks = w.K.data() + s * N;
        for (const auto& t : M.jac_terms)
        {
          const auto& rx = M.rx[t.process];
          const size_t ndp = t.dep_reactants.size();
          double dd = k[t.process] * dt[t.process];
          for (auto q : t.dep_reactants)
            dd *= y[q];
          for (size_t a = 0; a < ndp; ++a)
          {
            double part = k[t.process];
            for (size_t b = 0; b < ndp; ++b)
              if (b != a)
                part *= y[t.dep_reactants[b]];
            dd += part * dyd[t.dep_reactants[a]];
          }
          const double c = dd * Ks[t.independent];
          for (auto q : t.dep_reactants)
            dr[q] -= c;
          dr[t.independent] -= c;
          for (size_t j = 0; j < rx.products.size(); ++j)
            dr[rx.products[j]] += rx.yields[j] * c;
        }
        lu_solve(M, w.L.data(), w.U.data(), dr.data());
        std::memcpy(dK.data() + s * N, dr.data(), sizeof(double) * N);
      }
      for (int64_t i = 0; i < N; ++i)
      {
        double v = dyd[i];
        for (int s = 0; s < S; ++s)
          v += p.m[s] * dK[s * N + i];
        dyd[i] = v;
      }
    }
    std::memcpy(y, ynew.data(), sizeof(double) * N);
  }

  struct OracleSolver
  {
    Mech M;
  };
}  // namespace

extern "C"
{
  const char* oracle_last_error()
  {
    return g_error.c_str();
  }

  void* oracle_solver_create(const micm_hip_mechanism_desc* d, int32_t tableau)
  {
    try
    {
      auto s = std::make_unique<OracleSolver>();
      build(s->M, d, tableau);
      return s.release();
    }
    catch (const std::exception& e)
    {
      g_error = e.what();
      return nullptr;
    }
  }

  void oracle_solver_destroy(void* h)
  {
    delete static_cast<OracleSolver*>(h);
  }

  int oracle_sizes(void* h, int64_t* out)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
    out[0] = M.N;
    out[1] = M.R;
    out[2] = M.P;
    out[3] = M.nnzA;
    out[4] = M.nnzL;
    out[5] = M.nnzU;
    return 0;
  }

  int oracle_species_order(void* h, int64_t* state_to_declared)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
    for (int64_t i = 0; i < M.N; ++i)
      state_to_declared[i] = M.state_to_declared[i];
    return 0;
  }

  int oracle_patterns(void* h, int64_t* jr, int64_t* jc, int64_t* lr, int64_t* lc, int64_t* ur, int64_t* uc)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
    const int64_t n = M.N;
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < n; ++j)
      {
        if (M.a_pos[i * n + j] >= 0)
        {
          jr[M.a_pos[i * n + j]] = i;
          jc[M.a_pos[i * n + j]] = j;
        }
        if (M.l_pos[i * n + j] >= 0)
        {
          lr[M.l_pos[i * n + j]] = i;
          lc[M.l_pos[i * n + j]] = j;
        }
        if (M.u_pos[i * n + j] >= 0)
        {
          ur[M.u_pos[i * n + j]] = i;
          uc[M.u_pos[i * n + j]] = j;
        }
      }
    return 0;
  }

  int oracle_rate_constants(void* h, int64_t n_cells, const double* conditions, const double* custom, double* k)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cells; ++c)
      for (int64_t r = 0; r < M.R; ++r)
        k[c * M.R + r] = rate_constant(M.rx[r], conditions[3 * c], conditions[3 * c + 1], conditions[3 * c + 2],
                                       custom + c * M.P);
    return 0;
  }

  int oracle_forcing(void* h, int64_t n_cells, const double* k, const double* y, double* f)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cells; ++c)
    {
      for (int64_t i = 0; i < M.N; ++i)
        f[c * M.N + i] = 0.0;
      add_forcing(M, k + c * M.R, y + c * M.N, f + c * M.N);
    }
    return 0;
  }

  int oracle_jacobian(void* h, int64_t n_cells, const double* k, const double* y, double* jac)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cells; ++c)
    {
      for (int64_t i = 0; i < M.nnzA; ++i)
        jac[c * M.nnzA + i] = 0.0;
      subtract_jacobian(M, k + c * M.R, y + c * M.N, jac + c * M.nnzA);
    }
    return 0;
  }

  // AddForcingTerms / SubtractJacobianTerms accumulating into f / jac as they hold
  // (process_set.hpp:524-670), without the Fill(0) the two functions above apply first
  int oracle_add_forcing(void* h, int64_t n_cells, const double* k, const double* y, double* f)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cells; ++c)
      add_forcing(M, k + c * M.R, y + c * M.N, f + c * M.N);
    return 0;
  }

  int oracle_subtract_jacobian(void* h, int64_t n_cells, const double* k, const double* y, double* jac)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cells; ++c)
      subtract_jacobian(M, k + c * M.R, y + c * M.N, jac + c * M.nnzA);
    return 0;
  }

  int oracle_lu(void* h, int64_t n_cells, const double* jac, double alpha, double* jac_out, double* lower, double* upper)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cells; ++c)
    {
      double* A = jac_out + c * M.nnzA;
      std::memcpy(A, jac + c * M.nnzA, sizeof(double) * M.nnzA);
      for (int64_t i = 0; i < M.N; ++i)
        A[M.diag[i]] += alpha * 1.0;
      lu_decompose(M, A, lower + c * M.nnzL, upper + c * M.nnzU);
    }
    return 0;
  }

  int oracle_lu_solve(void* h, int64_t n_cells, const double* lower, const double* upper, double* x)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cells; ++c)
      lu_solve(M, lower + c * M.nnzL, upper + c * M.nnzU, x + c * M.N);
    return 0;
  }

  int oracle_normalized_error(void* h, int64_t n_cells, const double* y, const double* ynew, const double* yerr,
                              double rtol, const double* atol, double* error)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
    double e = 0;
    for (int64_t c = 0; c < n_cells; ++c)
      for (int64_t i = 0; i < M.N; ++i)
      {
        double a = y[c * M.N + i], b = ynew[c * M.N + i];
        double ymax = (std::abs(a) > std::abs(b) ? std::abs(a) : std::abs(b));
        double q = yerr[c * M.N + i] / (atol[i] + rtol * ymax);
        e += q * q;
      }
    const double Ntot = static_cast<double>(std::max<int64_t>(1, n_cells * M.N));
    *error = std::max(std::sqrt(e / Ntot), 1.0e-10);
    return 0;
  }

  // The same in the summation order of VectorMatrix<group> (group 1: Matrix<Real>).
  int oracle_normalized_error_grouped(void* h, int64_t n_cells, const double* y, const double* ynew,
                                      const double* yerr, double rtol, const double* atol, int64_t group,
                                      double* error)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
    double e = 0;
    for (int64_t g0 = 0; g0 < n_cells; g0 += group)
    {
      const int64_t rows = std::min(group, n_cells - g0);
      for (int64_t i = 0; i < M.N; ++i)
        for (int64_t c = g0; c < g0 + rows; ++c)
        {
          double a = y[c * M.N + i], b = ynew[c * M.N + i];
          double ymax = (std::abs(a) > std::abs(b) ? std::abs(a) : std::abs(b));
          double q = yerr[c * M.N + i] / (atol[i] + rtol * ymax);
          e += q * q;
        }
    }
    const double Ntot = static_cast<double>(std::max<int64_t>(1, n_cells * M.N));
    *error = std::max(std::sqrt(e / Ntot), 1.0e-10);
    return 0;
  }

  // lockstep: 0 one 1-cell State per cell; 1 one State of every cell (Matrix<Real> order);
  // L > 1 one State of every cell summing its error norm in VectorMatrix<L> order.
  int oracle_solve(void* h, int64_t n_cells, int32_t lockstep, const double* conditions, const double* custom, double* y,
                   double dt, int64_t n_solves, double rtol, const double* atol, int32_t* states, int64_t* stats,
                   double* final_time)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
    const double* at = atol ? atol : M.atol.data();
    if (rtol <= 0)
      rtol = 1e-6;
    auto run = [&](int64_t c0, int64_t nc)
    {
      std::vector<double> k(nc * M.R);
      int64_t acc[7] = {};
      Stats st;
      for (int64_t rep = 0; rep < n_solves; ++rep)
      {
        for (int64_t c = 0; c < nc; ++c)
          for (int64_t r = 0; r < M.R; ++r)
          {
            const double* cd = conditions + 3 * (c0 + c);
            k[c * M.R + r] = rate_constant(M.rx[r], cd[0], cd[1], cd[2], custom + (c0 + c) * M.P);
          }
        solve_state(M, nc, k.data(), y + c0 * M.N, dt, rtol, at, st, nullptr, lockstep > 1 ? lockstep : 1);
        for (int j = 0; j < 7; ++j)
          acc[j] += st.v[j];
      }
      for (int64_t c = 0; c < nc; ++c)
      {
        if (states)
          states[c0 + c] = st.state;
        if (stats)
          for (int j = 0; j < 7; ++j)
            stats[(c0 + c) * 7 + j] = acc[j];
        if (final_time)
          final_time[c0 + c] = st.final_time;
      }
    };
    try
    {
      if (lockstep)
        run(0, n_cells);
      else
      {
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t c = 0; c < n_cells; ++c)
          run(c, 1);
      }
    }
    catch (const std::exception& e)
    {
      g_error = e.what();
      return -1;
    }
    return 0;
  }

  // ---------------------------------------------------------------- discrete adjoint
  // Tapes: per cell max_records records of N + 1 doubles (see solve_state), counts[c]
  // records used; a count above max_records marks a tape that overflowed.

  // n_solves x Solve(dt) per cell from given rate constants, recording each cell's tape.
  int oracle_solve_record(void* h, int64_t n_cells, const double* k, double* y, double dt, int64_t n_solves, double rtol,
                          const double* atol, int32_t* states, int64_t* stats, int64_t max_records, double* tape,
                          int64_t* counts)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
    const double* at = atol ? atol : M.atol.data();
    if (rtol <= 0)
      rtol = 1e-6;
    const int64_t rec = M.N + 1;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t c = 0; c < n_cells; ++c)
    {
      std::vector<double> t;
      Stats st;
      int64_t acc[7] = {};
      for (int64_t rep = 0; rep < n_solves; ++rep)
      {
        solve_state(M, 1, k + c * M.R, y + c * M.N, dt, rtol, at, st, &t);
        for (int j = 0; j < 7; ++j)
          acc[j] += st.v[j];
      }
      const int64_t n = static_cast<int64_t>(t.size()) / rec;
      counts[c] = n;
      std::memcpy(tape + c * max_records * rec, t.data(), sizeof(double) * std::min(n, max_records) * rec);
      if (states)
        states[c] = st.state;
      if (stats)
        for (int j = 0; j < 7; ++j)
          stats[c * 7 + j] = acc[j];
    }
    return 0;
  }

  // The recorded map applied to y (in: initial, out: final): every step with its
  // recorded H, every clamp; the tapes' own states are not read.
  int oracle_replay(void* h, int64_t n_cells, const double* k, double* y, int64_t max_records, const double* tape,
                    const int64_t* counts)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
    const int64_t N = M.N, rec = N + 1;
    for (int64_t c = 0; c < n_cells; ++c)
      if (counts[c] > max_records)
      {
        g_error = "tape of cell " + std::to_string(c) + " overflowed";
        return -1;
      }
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t c = 0; c < n_cells; ++c)
    {
      StepData w;
      step_data_init(M, w);
      std::vector<double> ynew(N);
      double* yc = y + c * N;
      for (int64_t r = 0; r < counts[c]; ++r)
      {
        const double* rc = tape + (c * max_records + r) * rec;
        if (rc[0] == kClampTag)
          for (int64_t i = 0; i < N; ++i)
            yc[i] = std::max(yc[i], 0.0);
        else
        {
          frozen_step(M, k + c * M.R, yc, rc[0], w, ynew.data());
          std::memcpy(yc, ynew.data(), sizeof(double) * N);
        }
      }
    }
    return 0;
  }

  // Reverse sweep over the tapes: lam_final = dL/dy_final [cells][N] -> grad_y0 [cells][N]
  // and grad_logk [cells][R] = k dL/dk.
  int oracle_adjoint(void* h, int64_t n_cells, const double* k, int64_t max_records, const double* tape,
                     const int64_t* counts, const double* lam_final, double* grad_y0, double* grad_logk)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
    const int64_t N = M.N, rec = N + 1;
    for (int64_t c = 0; c < n_cells; ++c)
      if (counts[c] > max_records)
      {
        g_error = "tape of cell " + std::to_string(c) + " overflowed";
        return -1;
      }
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t c = 0; c < n_cells; ++c)
    {
      StepData w;
      step_data_init(M, w);
      double* lam = grad_y0 + c * N;
      double* th = grad_logk + c * M.R;
      std::memcpy(lam, lam_final + c * N, sizeof(double) * N);
      std::fill(th, th + M.R, 0.0);
      for (int64_t r = counts[c] - 1; r >= 0; --r)
      {
        const double* rc = tape + (c * max_records + r) * rec;
        if (rc[0] == kClampTag)
        {
          for (int64_t i = 0; i < N; ++i)
            if (rc[1 + i] < 0.0)
              lam[i] = 0.0;
        }
        else
          step_adjoint(M, k + c * M.R, rc + 1, rc[0], w, lam, th);
      }
    }
    return 0;
  }

  // Tangent-linear model of the recorded map along n_dir directions per cell:
  // dy0 [cells][n_dir][N], dlogk [cells][n_dir][R] -> dy_final [cells][n_dir][N]; y0 [cells][N].
  int oracle_tangent(void* h, int64_t n_cells, const double* k, const double* y0, int64_t max_records,
                     const double* tape, const int64_t* counts, int64_t n_dir, const double* dy0, const double* dlogk,
                     double* dy_final)
  {
    const Mech& M = static_cast<OracleSolver*>(h)->M;
    const int64_t N = M.N, rec = N + 1;
    for (int64_t c = 0; c < n_cells; ++c)
      if (counts[c] > max_records)
      {
        g_error = "tape of cell " + std::to_string(c) + " overflowed";
        return -1;
      }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < n_cells; ++c)
    {
      StepData w;
      step_data_init(M, w);
      std::vector<double> y(y0 + c * N, y0 + (c + 1) * N);
      double* dy = dy_final + c * n_dir * N;
      std::memcpy(dy, dy0 + c * n_dir * N, sizeof(double) * n_dir * N);
      for (int64_t r = 0; r < counts[c]; ++r)
      {
        const double* rc = tape + (c * max_records + r) * rec;
        if (rc[0] == kClampTag)
        {
          for (int64_t i = 0; i < N; ++i)
            if (y[i] < 0.0)
            {
              y[i] = 0.0;
              for (int64_t d = 0; d < n_dir; ++d)
                dy[d * N + i] = 0.0;
            }
        }
        else
          step_tangent(M, k + c * M.R, y.data(), rc[0], w, n_dir, dy, dlogk + c * n_dir * M.R);
      }
    }
    return 0;
  }
}

// Step limit of the solver's Solves (max_number_of_steps, rosenbrock_solver_parameters.hpp:31),
// raised for tight-tolerance accuracy references.
extern "C" int oracle_set_max_steps(void* h, int64_t n)
{
  if (!h || n < 1)
    return -1;
  static_cast<OracleSolver*>(h)->M.tab.max_steps = n;
  return 0;
}

// Step-control fields of the solver's parameters (rosenbrock_solver_parameters.hpp:31-41):
// max_number_of_steps, h_min, h_max, h_start; tests force the solver's failure states with them.
extern "C" int oracle_set_control(void* h, int64_t max_steps, double h_min, double h_max, double h_start)
{
  if (!h || max_steps < 0)
    return -1;
  Tableau& t = static_cast<OracleSolver*>(h)->M.tab;
  t.max_steps = max_steps;
  t.h_min = h_min;
  t.h_max = h_max;
  t.h_start = h_start;
  return 0;
}

// Plants a mistake in the adjoint (tests of the checks only; 0 restores it).
extern "C" int oracle_set_adjoint_fault(int32_t which)
{
  if (which < 0 || which > 3)
    return -1;
  g_adjoint_fault = which;
  return 0;
}

// Worker threads of the per-cell solves (OpenMP).
extern "C" int oracle_set_threads(int32_t n)
{
  if (n < 1)
    return -1;
  omp_set_num_threads(n);
  return 0;
}

#ifdef ORACLE_PORTABLE_MATH
// Element-wise evaluation of the portable math functions, for their accuracy
// tests: which = 0 exp(x), 1 log(x), 2 log10(x), 3 pow(x, y).
extern "C" int oracle_pmath(int32_t which, int64_t n, const double* x, const double* y, double* out)
{
  for (int64_t i = 0; i < n; ++i)
  {
    switch (which)
    {
      case 0: out[i] = pmath::exp(x[i]); break;
      case 1: out[i] = pmath::log(x[i]); break;
      case 2: out[i] = pmath::log10(x[i]); break;
      case 3: out[i] = pmath::pow(x[i], y[i]); break;
      default: return -1;
    }
  }
  return 0;
}
#endif
