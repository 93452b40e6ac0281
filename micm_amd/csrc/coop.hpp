// coop.hpp -- the cooperative form: one wavefront integrates one cell with the
// cell's whole working set (in-place LU, Jacobian, stage vectors) in LDS.
//
// Work inside a cell is spread over the 64 lanes without changing any rounding:
//   forcing   per-species gather, terms in the reference's scatter order
//             (process_set.hpp:541-579)
//   -J        per-element gather, terms in the reference's order (process_set.hpp:613-659)
//   LU        right-looking elimination over the Doolittle fill pattern; element
//             (k, m) receives its products in ascending j exactly as in
//             lu_decomposition_doolittle.inl:347-398, and L[k][j] is divided by U[j][j]
//             before it is used
//   forward   column-oriented; x_i receives L_ij x_j in ascending j (linear_solver.inl:250-265)
//   backward  row-oriented by one lane, ascending j, then the division (linear_solver.inl:270-287)
//   error     squares formed by all lanes, summed in species order by one lane
//             (rosenbrock.inl:377-392)
#pragma once

#include "device.hpp"
#include "path.hpp"

namespace micm_hip
{
  constexpr int kCoopThreads = 64;

  // Synchronisation inside the cooperative workgroup, which is exactly one wavefront:
  // its lanes issue in lockstep, so it is enough that this wave's LDS operations have
  // completed (lgkmcnt(0)) and that the compiler moves no memory access across the
  // point.  Unlike __syncthreads() it does not wait for outstanding global loads, so
  // prefetched table reads stay in flight.
  __device__ __forceinline__ void coop_sync()
  {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // vmcnt(63) expcnt(7) lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }

  struct CoopLds
  {
    double *alu, *jac, *y, *ynew, *f0, *yerr, *K, *k, *scratch;
  };

  __host__ __device__ inline size_t coop_lds_doubles(int nnz_alu, int nnz_jac, int n, int stages, int n_reactions,
                                                      int n_terms)
  {
    int scratch = n_reactions > n_terms ? n_reactions : n_terms;
    scratch = scratch > n ? scratch : n;
    return static_cast<size_t>(nnz_alu) + nnz_jac + static_cast<size_t>(4 + stages) * n + n_reactions + scratch;
  }

  // Bytes of the packed substitution tables staged after the per-cell data.
  __host__ __device__ inline size_t coop_table_bytes(int n_fcol, int n_brow)
  {
    return (static_cast<size_t>(n_fcol) + n_brow) * sizeof(uint32_t);
  }

  __device__ inline CoopLds coop_layout(double* lds, const DevMech& m, const DevCoop& c, int stages)
  {
    CoopLds L;
    const int n = m.n_species;
    L.alu = lds;
    L.jac = L.alu + c.nnz_alu;
    L.y = L.jac + m.nnz_jac;
    L.ynew = L.y + n;
    L.f0 = L.ynew + n;
    L.yerr = L.f0 + n;
    L.K = L.yerr + n;
    L.k = L.K + static_cast<size_t>(stages) * n;
    L.scratch = L.k + m.n_reactions;
    return L;
  }

  // out[s] = f(yv)[s]
  __device__ inline void coop_forcing(const DevMech& m, const DevCoop& c, const CoopLds& L, const double* yv,
                                      double* out)
  {
    const int lane = threadIdx.x;
    for (int r = lane; r < m.n_reactions; r += kCoopThreads)
    {
      double rate = L.k[r];
      for (int i = m.reactant_offsets[r]; i < m.reactant_offsets[r + 1]; ++i)
        rate *= yv[m.reactant_ids[i]];
      L.scratch[r] = rate;
    }
    coop_sync();
    for (int s = lane; s < m.n_species; s += kCoopThreads)
    {
      double v = 0.0;
      for (int q = c.spec_offsets[s]; q < c.spec_offsets[s + 1]; ++q)
        v += c.spec_coef[q] * L.scratch[c.spec_rxn[q]];
      out[s] = v;
    }
    coop_sync();
  }

  // L.jac = -J(y)
  __device__ inline void coop_jacobian(const DevMech& m, const DevCoop& c, const CoopLds& L)
  {
    const int lane = threadIdx.x;
    for (int t = lane; t < c.n_terms; t += kCoopThreads)
    {
      double d = L.k[m.term_process[t]];
      for (int i = m.term_dep_offsets[t]; i < m.term_dep_offsets[t + 1]; ++i)
        d *= L.y[m.term_dep_ids[i]];
      L.scratch[t] = d;
    }
    coop_sync();
    for (int e = lane; e < m.nnz_jac; e += kCoopThreads)
    {
      double v = 0.0;
      for (int q = c.jel_offsets[e]; q < c.jel_offsets[e + 1]; ++q)
        v += c.jel_coef[q] * L.scratch[c.jel_term[q]];
      L.jac[e] = v;
    }
    coop_sync();
  }

  // In-place LU of L.jac into L.alu.  Each lane applies the triples q = lane, lane+64, ...
  // of a step; a lane's first eight triples of step j + 1 are loaded while step j runs.
  __device__ inline void coop_decompose(const DevMech& m, const DevCoop& c, const CoopLds& L)
  {
    const int lane = threadIdx.x;
    for (int q = lane; q < c.nnz_alu; q += kCoopThreads)
      L.alu[q] = 0.0;
    coop_sync();
    for (int e = lane; e < m.nnz_jac; e += kCoopThreads)
      L.alu[c.jac_to_alu[e]] = L.jac[e];
    coop_sync();
    constexpr int B = 8;
    uint2 next[B];
    {
      const int b = c.upd_offsets[0] + lane, e = c.upd_offsets[1];
#pragma unroll
      for (int r = 0; r < B; ++r)
        if (b + r * kCoopThreads < e)
          next[r] = c.upd16[b + r * kCoopThreads];
    }
    for (int j = 0; j < m.n_species; ++j)
    {
      uint2 cur[B];
#pragma unroll
      for (int r = 0; r < B; ++r)
        cur[r] = next[r];
      if (j + 1 < m.n_species)
      {
        const int b = c.upd_offsets[j + 1] + lane, e = c.upd_offsets[j + 2];
#pragma unroll
        for (int r = 0; r < B; ++r)
          if (b + r * kCoopThreads < e)
            next[r] = c.upd16[b + r * kCoopThreads];
      }
      const double piv = L.alu[c.alu_diag[j]];
      for (int q = c.div_offsets[j] + lane; q < c.div_offsets[j + 1]; q += kCoopThreads)
        L.alu[c.div[q]] = L.alu[c.div[q]] / piv;
      coop_sync();
      // the triples of one step touch distinct targets, so their order is free
      const int b = c.upd_offsets[j] + lane, e = c.upd_offsets[j + 1];
#pragma unroll
      for (int r = 0; r < B; ++r)
        if (b + r * kCoopThreads < e)
        {
          const uint2 t = cur[r];
          L.alu[t.x & 0xffffu] -= L.alu[t.x >> 16] * L.alu[t.y];
        }
      for (int q0 = b + B * kCoopThreads; q0 < e; q0 += B * kCoopThreads)
      {
        uint2 tr[B];
#pragma unroll
        for (int r = 0; r < B; ++r)
          if (q0 + r * kCoopThreads < e)
            tr[r] = c.upd16[q0 + r * kCoopThreads];
#pragma unroll
        for (int r = 0; r < B; ++r)
          if (q0 + r * kCoopThreads < e)
            L.alu[tr[r].x & 0xffffu] -= L.alu[tr[r].x >> 16] * L.alu[tr[r].y];
      }
      coop_sync();
    }
  }

  // Substitution tables: packed 16-bit pairs, staged in LDS once per workgroup
  struct CoopSolveTables
  {
    const uint32_t* fcol;
    const uint32_t* brow;
  };

  __device__ inline void coop_solve(const DevMech& m, const DevCoop& c, const CoopLds& L, const CoopSolveTables& T,
                                    double* x)
  {
    const int lane = threadIdx.x;
    const int n = m.n_species;
    for (int j = 0; j < n; ++j)
    {
      const double xj = x[j];
      for (int q = c.fcol_offsets[j] + lane; q < c.fcol_offsets[j + 1]; q += kCoopThreads)
      {
        const uint32_t ip = T.fcol[q];
        x[ip & 0xffffu] -= L.alu[ip >> 16] * xj;
      }
      coop_sync();
    }
    if (lane == 0)
    {
      // row i of U right of the diagonal is contiguous in ALU; only the columns are looked up
      for (int i = n - 1; i >= 0; --i)
      {
        double v = x[i];
        const int d = c.alu_diag[i];
        int q = c.brow_offsets[i];
        const int e = c.brow_offsets[i + 1];
        const double* u = L.alu + d + 1 - q;
        constexpr int B = 16;
        for (; q + B <= e; q += B)
        {
          double a[B], b[B];
#pragma unroll
          for (int r = 0; r < B; ++r)
          {
            a[r] = u[q + r];
            b[r] = x[T.brow[q + r] >> 16];
          }
#pragma unroll
          for (int r = 0; r < B; ++r)
            v -= a[r] * b[r];
        }
        for (; q < e; ++q)
          v -= u[q] * x[T.brow[q] >> 16];
        x[i] = v / L.alu[d];
      }
    }
    coop_sync();
  }
}  // namespace micm_hip
