// adjoint.hpp -- the discrete adjoint of the per-cell Rosenbrock Solve.
//
// The reference has no adjoint.  What is differentiated is the map a recorded Solve
// applied to each cell: its accepted steps, each with its step size H frozen, and the
// post-solve clamp (solver.hpp:275-297) at the end of every Solve.  One accepted step
// (rosenbrock.inl:140-207, separate-LU semantics) with A = alpha I - J(y),
// alpha = 1 / (H gamma_0), f-evaluations g_s = f(Y_s) at Y_0 = y and
// Y_s = y + sum_{j<s} a_sj K_j where new_function_evaluation[s] (else g_s = g_{s-1}):
//
//     r_s = g_s + sum_{j<s} (c_sj / H) K_j,    K_s = A^-1 r_s,    ynew = y + sum_s m_s K_s.
//
// Reverse, lam = dL/dynew on entry:  Kbar_s = m_s lam, then for s = S-1 .. 0
//
//     rbar  = A^-T Kbar_s                       (U^T then L^T on the step's factors)
//     lam  += d/dy (rbar^T J(y) K_s)            (second derivative of the mass-action rates)
//     thb  += k d/dk (rbar^T J(y) K_s)
//     gbar += rbar,   Kbar_j += (c_sj / H) rbar
//     at an f-evaluation:  ybar = J(Y_s)^T gbar,  thb += k d/dk (gbar^T f(Y_s)),
//                          lam += ybar,  Kbar_j += a_sj ybar,  gbar = 0
//
// and lam = dL/dy on exit.  Both derivative terms go through one scalar per reaction,
// w_r(u) = sum_products yield u - sum_reactants u: with rate_r = k_r prod y_reactants,
// u^T f = sum_r w_r(u) rate_r and u^T J v = sum_t w_{r(t)}(u) d_t v_{q(t)} over the
// Jacobian terms t (reaction r(t), independent species q(t), d_t = k prod of the other
// reactants).  thb is d/dlog k per cell: for every rate type this is the derivative with
// respect to the log of a common factor on the reaction's A-prefactors.
//
// One workgroup integrates one cell: each record is replayed from its checkpoint (the
// Jacobian, the factorisation and the stage vectors are recomputed with the cooperative
// form's level schedules, so the stage vectors are those of the recorded step), then the
// stages are reversed.  Nothing per step is stored but the checkpoint.
#pragma once

#include "coop.hpp"

namespace micm_hip
{
  // Tables of the adjoint (MechanismTables: lv_ut, lv_lt, term_indep, ...).
  struct DevAdjoint
  {
    DevLevels ut, lt;
    const int32_t* term_indep;
    const int32_t* rxn_term_offsets;
    const int32_t* rxn_terms;
    const int32_t* indep_term_offsets;
    const int32_t* indep_terms;
    const int32_t* hes_offsets;
    const int32_t* hes_term;
    const int32_t* hes_slot;
  };

  // Checkpoints of recorded Solves: per cell max_records records of n_species + 1 doubles,
  // (H, y at the start of an accepted step) or (kClampTag, y before the post-solve clamp);
  // counts[cell] records written so far (above max_records: the tape overflowed and later
  // records were dropped).
  struct Tape
  {
    double* records;
    int64_t* counts;
    int64_t max_records;
  };
  constexpr double kClampTag = -1.0;

  __host__ __device__ inline size_t adjoint_lds_doubles(int nnz_alu, int n, int stages, int n_reactions, int n_terms)
  {
    const int tmp = n_reactions > n_terms ? n_reactions : n_terms;
    return static_cast<size_t>(nnz_alu) + static_cast<size_t>(5 + 2 * stages) * n + 3 * static_cast<size_t>(n_reactions) +
           n_terms + tmp;
  }

  // sum_products yield u[product] - sum_reactants u[reactant]
  __device__ __forceinline__ double reaction_weight(const DevMech& m, int r, const double* u)
  {
    double w = 0.0;
    for (int i = m.product_offsets[r]; i < m.product_offsets[r + 1]; ++i)
      w += m.yields[i] * u[m.product_ids[i]];
    for (int i = m.reactant_offsets[r]; i < m.reactant_offsets[r + 1]; ++i)
      w -= u[m.reactant_ids[i]];
    return w;
  }

  // d_t = k_r prod of the term's other reactants, optionally without dependent slot `skip`
  __device__ __forceinline__ double term_product(const DevMech& m, int t, const double* k, const double* y,
                                                 int skip = -1)
  {
    double d = k[m.term_process[t]];
    const int b = m.term_dep_offsets[t], e = m.term_dep_offsets[t + 1];
    for (int i = b; i < e; ++i)
      if (i - b != skip)
        d *= y[m.term_dep_ids[i]];
    return d;
  }

  template<int NT>
  __global__ void __launch_bounds__(NT) k_adjoint(DevMech m, DevCoop c, DevAdjoint a, DevParams p, int64_t n_cells,
                                                  const double* __restrict__ k_state, Tape tape,
                                                  const double* __restrict__ seed, double* __restrict__ grad_y,
                                                  double* __restrict__ grad_logk)
  {
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    const int n = m.n_species, R = m.n_reactions, S = p.stages, nt = c.n_terms;
    double* alu = lds;
    double* y = alu + c.nnz_alu;
    double* ys = y + n;
    double* lam = ys + n;
    double* gb = lam + n;
    double* rb = gb + n;
    double* K = rb + n;
    double* Kb = K + static_cast<size_t>(S) * n;
    double* kk = Kb + static_cast<size_t>(S) * n;
    double* w = kk + R;
    double* th = w + R;
    double* jterm = th + R;
    double* tmp = jterm + nt;
    CoopLds L{};
    L.alu = alu;
    L.y = y;
    L.k = kk;
    L.scratch = tmp;
    L.jterm = jterm;
    const size_t rec_len = static_cast<size_t>(n) + 1;
    // the forward substitution's levels before its column sweep, then the sweep (whose
    // reads past a row's last factor entry stay inside this layout's LDS, unused)
    const bool sweep = c.fwd_sweep.first_level >= 0;
    const uint4 me = sweep && tid < kWave ? c.fwd_sweep.lanes[tid] : make_uint4(0u, 0u, 0u, 0u);
    DevLevels fwd = c.fwd;
    if (sweep)
      fwd.n_levels = c.fwd_sweep.first_level;
    for (int64_t cell = blockIdx.x; cell < n_cells; cell += gridDim.x)
    {
      const size_t g = cell / kGroup;
      const int cl = cell % kGroup;
      {
        Lane ks = lane_view(const_cast<double*>(k_state), g, R, cl);
        Lane sd = lane_view(const_cast<double*>(seed), g, n, cl);
        for (int r = tid; r < R; r += NT)
        {
          kk[r] = ks[r];
          th[r] = 0.0;
        }
        for (int i = tid; i < n; i += NT)
          lam[i] = sd[i];
      }
      const int64_t count = tape.counts[cell];
      if (count > tape.max_records)
      {
        // an overflowed tape lacks the cell's first records: no gradient
        Lane gy = lane_view(grad_y, g, n, cl);
        Lane gk = lane_view(grad_logk, g, R, cl);
        for (int i = tid; i < n; i += NT)
          gy[i] = __builtin_nan("");
        for (int r = tid; r < R; r += NT)
          gk[r] = __builtin_nan("");
        coop_sync<NT>();
        continue;
      }
      coop_sync<NT>();
      for (int64_t rix = count - 1; rix >= 0; --rix)
      {
        const double* rec = tape.records + (static_cast<size_t>(cell) * tape.max_records + rix) * rec_len;
        const double H = rec[0];
        if (H == kClampTag)
        {
          // y = max(y, 0): no derivative where the value was clamped
          for (int i = tid; i < n; i += NT)
            if (rec[1 + i] < 0.0)
              lam[i] = 0.0;
          coop_sync<NT>();
          continue;
        }
        // ---- the recorded step again: -J(y) into the factor, alpha, LU, stages
        for (int i = tid; i < n; i += NT)
          y[i] = rec[1 + i];
        for (int q = tid; q < c.nnz_alu; q += NT)
          alu[q] = 0.0;
        coop_sync<NT>();
        coop_jacobian<NT, true>(m, c, L);  // jterm = d_t(y)
        const double alpha = 1.0 / (H * p.gamma0);
        for (int i = tid; i < n; i += NT)
          alu[c.alu_diag[i]] += alpha * 1.0;
        coop_sync<NT>();
        // (the LU's records name the zero past the factor in their padded slots, which
        // this layout does not keep: the unpadded executor tests every slot)
        coop_levels_piped<NT, false>(c.lu, c.lu.terms, alu, alu, alu);
        for (int s = 0; s < S; ++s)
        {
          const int sc = s * (s - 1) / 2;
          double* Ks = K + static_cast<size_t>(s) * n;
          if (s == 0)
          {
            coop_forcing<NT>(m, c, L, y, Ks);
          }
          else if (p.new_f[s])
          {
            for (int i = tid; i < n; i += NT)
            {
              double v = y[i];
              for (int j = 0; j < s; ++j)
                v += p.a[sc + j] * K[static_cast<size_t>(j) * n + i];
              ys[i] = v;
            }
            coop_sync<NT>();
            coop_forcing<NT>(m, c, L, ys, Ks);
          }
          for (int i = tid; i < n; i += NT)
          {
            if (s + 1 < S && !p.new_f[s + 1])
              Ks[static_cast<size_t>(n) + i] = Ks[i];
            double v = Ks[i];
            for (int j = 0; j < s; ++j)
              v += (p.c[sc + j] / H) * K[static_cast<size_t>(j) * n + i];
            Ks[i] = v;
          }
          coop_sync<NT>();
          coop_levels_piped<NT, false>(fwd, c.fwd.terms, alu, Ks, Ks);
          if (sweep)
            coop_sweep<NT>(c.fwd_sweep, me, alu, Ks);
          coop_levels_piped<NT, false>(c.bwd, c.bwd.terms, alu, Ks, Ks);
        }
        // ---- reverse over the stages
        for (int i = tid; i < n; i += NT)
        {
          for (int s = 0; s < S; ++s)
            Kb[static_cast<size_t>(s) * n + i] = p.m[s] * lam[i];
          gb[i] = 0.0;
        }
        coop_sync<NT>();
        for (int s = S - 1; s >= 0; --s)
        {
          const int sc = s * (s - 1) / 2;
          const double* Ks = K + static_cast<size_t>(s) * n;
          for (int i = tid; i < n; i += NT)
            rb[i] = Kb[static_cast<size_t>(s) * n + i];
          coop_sync<NT>();
          coop_levels_piped<NT, false>(a.ut, a.ut.terms, alu, rb, rb);
          coop_levels_piped<NT, false>(a.lt, a.lt.terms, alu, rb, rb);
          for (int r = tid; r < R; r += NT)
            w[r] = reaction_weight(m, r, rb);
          coop_sync<NT>();
          // Jbar = rbar K_s^T through y and k
          for (int r = tid; r < R; r += NT)
          {
            double acc = 0.0;
            for (int q = a.rxn_term_offsets[r]; q < a.rxn_term_offsets[r + 1]; ++q)
            {
              const int t = a.rxn_terms[q];
              acc += jterm[t] * Ks[a.term_indep[t]];
            }
            th[r] += w[r] * acc;
          }
          for (int i = tid; i < n; i += NT)
          {
            double acc = 0.0;
            for (int q = a.hes_offsets[i]; q < a.hes_offsets[i + 1]; ++q)
            {
              const int t = a.hes_term[q];
              acc += w[m.term_process[t]] * Ks[a.term_indep[t]] * term_product(m, t, kk, y, a.hes_slot[q]);
            }
            lam[i] += acc;
            gb[i] += rb[i];
            for (int j = 0; j < s; ++j)
              Kb[static_cast<size_t>(j) * n + i] += (p.c[sc + j] / H) * rb[i];
          }
          coop_sync<NT>();
          if (s == 0 || p.new_f[s])
          {
            for (int i = tid; i < n; i += NT)
            {
              double v = y[i];
              for (int j = 0; j < s; ++j)
                v += p.a[sc + j] * K[static_cast<size_t>(j) * n + i];
              ys[i] = s == 0 ? y[i] : v;
            }
            for (int r = tid; r < R; r += NT)
              w[r] = reaction_weight(m, r, gb);
            coop_sync<NT>();
            for (int r = tid; r < R; r += NT)
            {
              double rate = kk[r];
              for (int i = m.reactant_offsets[r]; i < m.reactant_offsets[r + 1]; ++i)
                rate *= ys[m.reactant_ids[i]];
              th[r] += w[r] * rate;
            }
            for (int t = tid; t < nt; t += NT)
              tmp[t] = term_product(m, t, kk, ys);
            coop_sync<NT>();
            for (int i = tid; i < n; i += NT)
            {
              double yb = 0.0;
              for (int q = a.indep_term_offsets[i]; q < a.indep_term_offsets[i + 1]; ++q)
              {
                const int t = a.indep_terms[q];
                yb += w[m.term_process[t]] * tmp[t];
              }
              lam[i] += yb;
              for (int j = 0; j < s; ++j)
                Kb[static_cast<size_t>(j) * n + i] += p.a[sc + j] * yb;
              gb[i] = 0.0;
            }
            coop_sync<NT>();
          }
        }
      }
      {
        Lane gy = lane_view(grad_y, g, n, cl);
        Lane gk = lane_view(grad_logk, g, R, cl);
        for (int i = tid; i < n; i += NT)
          gy[i] = lam[i];
        for (int r = tid; r < R; r += NT)
          gk[r] = th[r];
      }
      coop_sync<NT>();
    }
  }
}  // namespace micm_hip
