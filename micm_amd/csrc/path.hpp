// path.hpp -- per-lane device functions of the Rosenbrock path.  Each
// reproduces the arithmetic order of the reference function it replaces, so
// that with FMA contraction disabled the results are bit-identical to the
// reference's CPU solver for +, -, *, / and sqrt; exp, log10 and pow come from
// pmath.h, whose results are the same on the host and on the device.
#pragma once

#include "device.hpp"
#include "pmath.h"

#include "../../include/micm_hip.h"

namespace micm_hip
{
  constexpr double kPi = 3.14159265358979323846;

  // std::min / std::max semantics (first argument returned on ties and NaN comparisons)
  __device__ __forceinline__ double smin(double a, double b)
  {
    return (b < a) ? b : a;
  }
  __device__ __forceinline__ double smax(double a, double b)
  {
    return (a < b) ? b : a;
  }

  // One rate constant (rate_constant_functions.hpp:51-136) times the
  // parameterized-reactant multipliers (reaction_rate_store.hpp:67-75, 510-517).
  __device__ inline double rate_constant(const DevMech& m, int r, double T, double P, double rho, const Lane& custom)
  {
    const double* p = m.rate_params + static_cast<size_t>(r) * MICM_HIP_RATE_PARAM_STRIDE;
    double k = 0.0;
    switch (m.rate_type[r])
    {
      case MICM_HIP_RATE_ARRHENIUS: k = p[0] * pmath::exp(p[2] / T) * pmath::pow(T / p[3], p[1]) * (1.0 + p[4] * P); break;
      case MICM_HIP_RATE_TROE:
      case MICM_HIP_RATE_TERNARY:
      {
        const double scale = m.rate_type[r] == MICM_HIP_RATE_TROE ? rho : 1.0;
        const double k0 = p[0] * pmath::exp(p[2] / T) * pmath::pow(T / 300.0, p[1]);
        const double kinf = p[3] * pmath::exp(p[5] / T) * pmath::pow(T / 300.0, p[4]);
        const double ratio = k0 * rho / kinf;
        // std::pow(x, 2.0) with the constant exponent 2 compiles to x * x (an exact folding
        // compilers make at any optimisation level), so the reference squares with a multiply
        const double lg = pmath::log10(ratio);
        k = k0 * scale / (1.0 + ratio) * pmath::pow(p[6], p[7] / (p[7] + lg * lg));
        break;
      }
      case MICM_HIP_RATE_BRANCHED:
      {
        const double a = p[5] * rho;
        const double b = 0.43 * pmath::pow(T / 298.0, -8.0);
        const double lg = pmath::log10(a / b);  // std::pow(lg, 2.0) compiles to lg * lg
        const double A_val = a / (1.0 + a / b) * pmath::pow(0.41, 1.0 / (1.0 + lg * lg));
        const double pre = p[1] * pmath::exp(-p[2] / T);
        k = p[0] == 0.0 ? pre * (p[6] / (p[6] + A_val)) : pre * (A_val / (A_val + p[6]));
        break;
      }
      case MICM_HIP_RATE_TUNNELING: k = p[0] * pmath::exp(-p[1] / T + p[2] / (T * T * T)); break;
      case MICM_HIP_RATE_TAYLOR:
      {
        double poly = 0.0, t_pow = 1.0;
        const int nc = static_cast<int>(p[5]);
        for (int j = 0; j < nc; ++j, t_pow *= T)
          poly += p[6 + j] * t_pow;
        k = poly * p[0] * pmath::exp(p[2] / T) * pmath::pow(T / p[3], p[1]) * (1.0 + p[4] * P);
        break;
      }
      case MICM_HIP_RATE_REVERSIBLE: k = p[0] * pmath::exp(p[1] / T) * p[2]; break;
      case MICM_HIP_RATE_USER_DEFINED: k = custom[m.custom_index[r]] * p[0]; break;
      case MICM_HIP_RATE_SURFACE:
      {
        const double radius = custom[m.custom_index[r]], num = custom[m.custom_index[r] + 1];
        const double mfs = sqrt(p[3] * T);
        k = 4.0 * num * kPi * radius * radius / (radius / p[0] + 4.0 / (mfs * p[2]));
        break;
      }
      default: break;
    }
    const int mb = m.multiplier_offsets[r], me = m.multiplier_offsets[r + 1];
    if (mb < me)
    {
      double v = 1.0;
      for (int i = mb; i < me; ++i)
      {
        const double* f = m.multipliers + 4 * i;
        v *= f[0] + f[1] * T + f[2] * P + f[3] * rho;
      }
      k *= v;
    }
    return k;
  }

  // forcing += f(y)  (process_set.hpp:524-590)
  __device__ inline void add_forcing(const DevMech& m, const Lane& k, const Lane& y, const Lane& f)
  {
    for (int r = 0; r < m.n_reactions; ++r)
    {
      double rate = k[r];
      const int rb = m.reactant_offsets[r], re = m.reactant_offsets[r + 1];
      for (int i = rb; i < re; ++i)
        rate *= y[m.reactant_ids[i]];
      for (int i = rb; i < re; ++i)
        f[m.reactant_ids[i]] -= rate;
      const int pb = m.product_offsets[r], pe = m.product_offsets[r + 1];
      for (int i = pb; i < pe; ++i)
        f[m.product_ids[i]] += m.yields[i] * rate;
    }
  }

  // jacobian -= J(y), i.e. accumulate -J  (process_set.hpp:594-670)
  __device__ inline void subtract_jacobian(const DevMech& m, const Lane& k, const Lane& y, const Lane& jac)
  {
    for (int t = 0; t < m.n_terms; ++t)
    {
      const int p = m.term_process[t];
      double d = k[p];
      const int db = m.term_dep_offsets[t], de = m.term_dep_offsets[t + 1];
      for (int i = db; i < de; ++i)
        d *= y[m.term_dep_ids[i]];
      int f = m.term_flat_offsets[t];
      for (int i = db; i < de; ++i)
        jac[m.term_flat[f++]] += d;
      jac[m.term_flat[f++]] += d;
      const int pb = m.product_offsets[p], pe = m.product_offsets[p + 1];
      for (int i = pb; i < pe; ++i)
        jac[m.term_flat[f++]] -= m.yields[i] * d;
    }
  }

  // v - sum_t L[pairs[t].x] * U[pairs[t].y] for t in [b, e), subtracting in order.  The
  // loads of eight terms are issued before their subtractions so that a lane keeps
  // sixteen memory requests in flight; the subtraction order is unchanged.
  __device__ __forceinline__ double subtract_products(double v, const int2* __restrict__ pairs, int b, int e,
                                                      const Lane& X, const Lane& Y)
  {
    int t = b;
    for (; t + 8 <= e; t += 8)
    {
      double x[8], y[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
      {
        const int2 ij = pairs[t + q];
        x[q] = X[ij.x];
        y[q] = Y[ij.y];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        v -= x[q] * y[q];
    }
    for (; t < e; ++t)
    {
      const int2 ij = pairs[t];
      v -= X[ij.x] * Y[ij.y];
    }
    return v;
  }

  // Doolittle LU of A into separate L (unit diagonal stored) and U
  // (lu_decomposition_doolittle.inl:325-403); sums run over j ascending.
  __device__ inline void lu_decompose(const DevMech& m, const Lane& A, const Lane& L, const Lane& U)
  {
    for (int i = 0; i < m.n_species; ++i)
      L[m.lower_diag[i]] = 1.0;
    for (int e = 0; e < m.n_lu_elements; ++e)
    {
      const int a = m.lu_a[e];
      double v = a >= 0 ? A[a] : 0.0;
      v = subtract_products(v, m.lu_terms, m.lu_term_offsets[e], m.lu_term_offsets[e + 1], L, U);
      const int div = m.lu_div[e];
      if (div < 0)
        U[m.lu_out[e]] = v;
      else
        L[m.lu_out[e]] = v / U[div];
    }
  }

  // x <- U^-1 L^-1 x  (linear_solver.inl:235-294).  With unit_lower the division
  // by L[i][i] is skipped, which is exact when L[i][i] == 1 (as lu_decompose writes).
  template<bool unit_lower>
  __device__ inline void lu_solve(const DevMech& m, const Lane& L, const Lane& U, const Lane& x)
  {
    const int n = m.n_species;
    for (int i = 0; i < n; ++i)
    {
      double v = subtract_products(x[i], m.fwd_terms, m.fwd_offsets[i], m.fwd_offsets[i + 1], L, x);
      if (!unit_lower)
        v /= L[m.lower_diag[i]];
      x[i] = v;
    }
    for (int i = n - 1; i >= 0; --i)
    {
      const double v = subtract_products(x[i], m.bwd_terms, m.bwd_offsets[i], m.bwd_offsets[i + 1], U, x);
      x[i] = v / U[m.upper_diag[i]];
    }
  }

  // sum_i (err_i / (atol_i + rtol * max(|y_i|, |ynew_i|)))^2 over species ascending
  // (rosenbrock.inl:377-392)
  __device__ inline double error_sum(int n, const Lane& y, const Lane& ynew, const Lane& err, const double* atol,
                                     double rtol)
  {
    // terms of eight species formed together (their loads in flight at once), then
    // added one after another in species order
    double acc = 0.0;
    int i = 0;
    for (; i + 8 <= n; i += 8)
    {
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
      {
        const double a = fabs(y[i + u]), b = fabs(ynew[i + u]);
        const double ymax = a > b ? a : b;
        const double q = err[i + u] / (atol[i + u] + rtol * ymax);
        t[u] = q * q;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        acc += t[u];
    }
    for (; i < n; ++i)
    {
      const double a = fabs(y[i]), b = fabs(ynew[i]);
      const double ymax = a > b ? a : b;
      const double q = err[i] / (atol[i] + rtol * ymax);
      acc += q * q;
    }
    return acc;
  }

  // ------------------------------------------------------------------ step control

  __device__ inline void control_init(const DevParams& p, CellControl& c)
  {
    const double H = smin(smax(p.h_min, fabs(p.h_start)), fabs(p.h_max));  // rosenbrock.inl:36
    c.H = H;
    c.t = 0.0;
    c.last_alpha = 0.0;
    c.flags = kNewStep;
    c.status = MICM_HIP_STATE_RUNNING;
    for (int i = 0; i < 7; ++i)
      c.stats[i] = 0;
  }

  // Loop-top checks of a new internal step (rosenbrock.inl:80-95).  Returns false
  // when the cell has finished (status set).
  __device__ inline bool begin_step(const DevParams& p, CellControl& c)
  {
    if (!((c.t - p.dt + p.round_off) <= 0))
    {
      c.status = MICM_HIP_STATE_CONVERGED;
      return false;
    }
    if (c.stats[MICM_HIP_STAT_STEPS] > p.max_steps)
    {
      c.status = MICM_HIP_STATE_MAX_STEPS;
      return false;
    }
    if (((c.t + 0.1 * c.H) == c.t) || (c.H <= p.round_off))
    {
      c.status = MICM_HIP_STATE_STEP_SIZE_TOO_SMALL;
      return false;
    }
    c.H = smin(c.H, fabs(p.dt - c.t));
    return true;
  }

  // Alpha increment of this attempt for the separate-LU path (rosenbrock.inl:125-134).
  __device__ inline double next_alpha(const DevParams& p, CellControl& c)
  {
    double alpha = 1.0 / (c.H * p.gamma0);
    const double cumulative = alpha;
    alpha -= c.last_alpha;
    c.last_alpha = cumulative;
    return alpha;
  }

  // Accept / reject with the normalized error (rosenbrock.inl:229-292).  Returns
  // true when the step was accepted (the caller then takes Ynew as Y).
  __device__ inline bool control_update(const DevParams& p, CellControl& c, double error)
  {
    const double fac = smin(p.factor_max, smax(p.factor_min, p.safety / pmath::pow(error, p.inv_elo)));
    double Hnew = c.H * fac;
    c.stats[MICM_HIP_STAT_STEPS] += 1;
    if (isnan(error))
    {
      c.status = MICM_HIP_STATE_NAN_DETECTED;
      return false;
    }
    if (isinf(error))
    {
      c.status = MICM_HIP_STATE_INF_DETECTED;
      return false;
    }
    if ((error < 1) || (c.H < p.h_min))
    {
      c.stats[MICM_HIP_STAT_ACCEPTED] += 1;
      c.t = c.t + c.H;
      Hnew = smax(p.h_min, smin(Hnew, p.h_max));
      if (c.flags & kRejectLast)
        Hnew = smin(Hnew, c.H);
      c.flags = kNewStep;
      c.H = Hnew;
      if (!((c.t - p.dt + p.round_off) <= 0))
        c.status = MICM_HIP_STATE_CONVERGED;
      return true;
    }
    if (c.flags & kRejectMore)
      Hnew = c.H * p.reject_decrease;
    int32_t flags = kRejectLast;
    if (c.flags & kRejectLast)
      flags |= kRejectMore;
    c.flags = flags;
    c.H = Hnew;
    if (c.stats[MICM_HIP_STAT_ACCEPTED] >= 1)
      c.stats[MICM_HIP_STAT_REJECTED] += 1;
    return false;
  }
}  // namespace micm_hip
