// coop.hpp -- the cooperative form: one workgroup (NT = 64 to 1024 threads)
// integrates one cell with the cell's working set (in-place LU, Jacobian, stage
// vectors) in LDS; the compact layout keeps only the gather sources there.
//
// Work inside a cell is spread over the threads without changing any rounding:
//   forcing   per-species gather, terms in the reference's scatter order
//             (process_set.hpp:541-579)
//   -J        per-element gather, terms in the reference's order (process_set.hpp:613-659)
//   LU        level-scheduled Doolittle: each element of the fill pattern subtracts
//             L[i][j] U[j][k] for j ascending and divides L elements by U[k][k], as
//             in lu_decomposition_doolittle.inl:347-398, its sum split into segments
//             that run as soon as their operands are final (the partial sum kept in
//             the element's position)
//   forward   level-scheduled rows, x_i -= L_ij x_j for j ascending (linear_solver.inl:250-265)
//   backward  level-scheduled rows, x_i -= U_ij x_j for j ascending, then / U_ii
//             (linear_solver.inl:270-287)
//   error     squares formed by all threads, summed in species order by one thread
//             (rosenbrock.inl:377-392)
#pragma once

#include "device.hpp"
#include "mechanism.hpp"
#include "path.hpp"

namespace micm_hip
{
  constexpr int kWave = 64;

  // Synchronisation inside the cooperative workgroup.  With a single wavefront its
  // lanes issue in lockstep, so it is enough that this wave's LDS operations have
  // completed (lgkmcnt(0)) and that the compiler moves no memory access across the
  // point; unlike __syncthreads() this does not wait for outstanding global loads.
  template<int NT>
  __device__ __forceinline__ void coop_sync()
  {
    if constexpr (NT == kWave)
    {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_s_waitcnt(0xc07f);  // vmcnt(63) expcnt(7) lgkmcnt(0)
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    else
    {
      __syncthreads();
    }
  }

  struct CoopLds
  {
    double *alu, *jac, *y, *ynew, *f0, *yerr, *K, *k, *scratch;
    double* jterm;     // Jacobian partial products: the ALU array when it is large enough
    double* products;  // per-level products of the LU and substitution schedules
    int p_capacity;    // doubles in products
  };

  // Doubles past the product buffer's capacity that the two-phase back substitution may
  // read (one batch of eight ahead of a task's last product, never used).
  constexpr int kProductSlack = 16;

  __host__ __device__ inline int coop_scratch_doubles(int n, int n_reactions, int nnz_alu, int n_terms)
  {
    int scratch = n_reactions > n ? n_reactions : n;
    if (n_terms > nnz_alu && n_terms > scratch)
      scratch = n_terms;
    return scratch;
  }

  // Doubles of per-cell LDS data, the per-level product buffer (p_capacity) included.
  // The Jacobian's partial products live in the ALU array (rebuilt from the Jacobian
  // before every factorisation) when it is large enough.
  __host__ __device__ inline size_t coop_lds_doubles(int nnz_alu, int nnz_jac, int n, int stages, int n_reactions,
                                                      int n_terms, int p_capacity)
  {
    return static_cast<size_t>(nnz_alu) + 1 + nnz_jac + static_cast<size_t>(4 + stages) * n + n_reactions +
           coop_scratch_doubles(n, n_reactions, nnz_alu, n_terms) + p_capacity + kProductSlack;
  }

  __device__ inline CoopLds coop_layout(double* lds, const DevMech& m, const DevCoop& c, int stages, int p_capacity)
  {
    CoopLds L;
    const int n = m.n_species;
    L.alu = lds;
    L.jac = L.alu + c.nnz_alu + 1;  // alu[nnz_alu] = 0 (see coop_decompose)
    L.y = L.jac + m.nnz_jac;
    L.ynew = L.y + n;
    L.f0 = L.ynew + n;
    L.yerr = L.f0 + n;
    L.K = L.yerr + n;
    L.k = L.K + static_cast<size_t>(stages) * n;
    L.scratch = L.k + m.n_reactions;
    L.jterm = c.n_terms <= c.nnz_alu ? L.alu : L.scratch;
    L.products = L.scratch + coop_scratch_doubles(n, m.n_reactions, c.nnz_alu, c.n_terms);
    L.p_capacity = p_capacity;
    return L;
  }

  // Compact layout: only what the gathers read stays in LDS -- the in-place factor, y
  // and the stage vectors (the current stage's forcing input and its solve vector) --
  // plus one scratch area shared in time by the reaction rates, the Jacobian partial
  // products (when they do not fit over the stage vectors), the back substitution's
  // products and the error terms.  The Jacobian is assembled straight into the factor
  // at every attempt (its diagonal, which accumulates the attempts' alpha shifts, and
  // f(y) are kept in registers, one species per thread) and the rate constants are
  // read from a per-workgroup copy in global memory.  About 20 % smaller than the
  // full layout, which fits a second cell per CU for mechanisms of 210/560's size.
  __host__ __device__ inline size_t coop_compact_scratch_doubles(int n, int n_reactions, int stages, int n_terms,
                                                                 int p_capacity)
  {
    int scratch = n_reactions > n ? n_reactions : n;
    if (p_capacity + kProductSlack > scratch)
      scratch = p_capacity + kProductSlack;
    if (n_terms > stages * n && n_terms > scratch)
      scratch = n_terms;
    return static_cast<size_t>(scratch);
  }
  __host__ __device__ inline size_t coop_compact_lds_doubles(int nnz_alu, int n, int stages, int n_reactions,
                                                             int n_terms, int p_capacity)
  {
    return static_cast<size_t>(nnz_alu) + 1 + static_cast<size_t>(1 + stages) * n +
           coop_compact_scratch_doubles(n, n_reactions, stages, n_terms, p_capacity);
  }
  __device__ inline CoopLds coop_layout_compact(double* lds, const DevMech& m, const DevCoop& c, int stages,
                                                int p_capacity, double* k_global)
  {
    CoopLds L;
    const int n = m.n_species;
    L.alu = lds;
    L.y = L.alu + c.nnz_alu + 1;  // alu[nnz_alu] = 0
    L.K = L.y + n;
    L.scratch = L.K + static_cast<size_t>(stages) * n;
    L.jac = nullptr;
    L.ynew = nullptr;
    L.f0 = nullptr;
    L.yerr = nullptr;
    L.k = k_global;
    L.jterm = c.n_terms <= stages * n ? L.K : L.scratch;
    L.products = L.scratch;
    L.p_capacity = p_capacity;
    return L;
  }

  // product of up to three packed operands, in list order, onto v; 0xffff counts
  // fall back to the index lists
  __device__ __forceinline__ double packed_product(double v, uint64_t pk, double y0, double y1, double y2)
  {
    const uint32_t count = static_cast<uint32_t>(pk >> 48);
    if (count >= 1)
      v *= y0;
    if (count >= 2)
      v *= y1;
    if (count >= 3)
      v *= y2;
    return v;
  }

  // v + sum_q coef[q] * src[idx[q] * TC + cc] for q in [b, e), added in order, in batches
  // of B <= 8 (the lists end in eight padding entries)
  template<int TC = 1, int B = 8>
  __device__ __forceinline__ double gather_sum(double v, const int32_t* idx, const double* coef, int b, int e,
                                               const double* src, int cc = 0)
  {
    static_assert(B >= 1 && B <= 8, "the lists end in eight padding entries");
    for (int q = b; q < e; q += B)
    {
      int ix[B];
      double cf[B], sv[B];
#pragma unroll
      for (int r = 0; r < B; ++r)
      {
        ix[r] = idx[q + r];
        cf[r] = coef[q + r];
      }
#pragma unroll
      for (int r = 0; r < B; ++r)
        sv[r] = src[ix[r] * TC + cc];
#pragma unroll
      for (int r = 0; r < B; ++r)
        if (q + r < e)
          v += cf[r] * sv[r];
    }
    return v;
  }

  // out[s] = f(yv)[s]
  template<int NT>
  __device__ inline void coop_forcing(const DevMech& m, const DevCoop& c, const CoopLds& L, const double* yv,
                                      double* out)
  {
    const int tid = threadIdx.x;
    const int R = m.n_reactions;
    for (int r0 = tid; r0 < R; r0 += 4 * NT)
    {
      uint64_t pk[4];
      double kk[4], y0[4], y1[4], y2[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        const int r = r0 + j * NT < R ? r0 + j * NT : 0;
        pk[j] = c.rxn_pack[r];
        kk[j] = L.k[r];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        y0[j] = yv[pk[j] & 0xffffu];
        y1[j] = yv[(pk[j] >> 16) & 0xffffu];
        y2[j] = yv[(pk[j] >> 32) & 0xffffu];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        const int r = r0 + j * NT;
        if (r < R)
        {
          double rate = kk[j];
          if ((pk[j] >> 48) == 0xffffu)
          {
            for (int i = m.reactant_offsets[r]; i < m.reactant_offsets[r + 1]; ++i)
              rate *= yv[m.reactant_ids[i]];
          }
          else
          {
            rate = packed_product(rate, pk[j], y0[j], y1[j], y2[j]);
          }
          L.scratch[r] = rate;
        }
      }
    }
    coop_sync<NT>();
    for (int s = tid; s < m.n_species; s += NT)
      out[s] = gather_sum(0.0, c.spec_rxn, c.spec_coef, c.spec_offsets[s], c.spec_offsets[s + 1], L.scratch);
    coop_sync<NT>();
  }

  // L.jac = -J(y); with TO_ALU the elements go straight to their factor positions
  template<int NT, bool TO_ALU = false>
  __device__ inline void coop_jacobian(const DevMech& m, const DevCoop& c, const CoopLds& L)
  {
    const int tid = threadIdx.x;
    const int nt = c.n_terms;
    for (int t0 = tid; t0 < nt; t0 += 4 * NT)
    {
      uint64_t pk[4];
      double kk[4], y0[4], y1[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        const int t = t0 + j * NT < nt ? t0 + j * NT : 0;
        pk[j] = c.term_pack[t];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        kk[j] = L.k[pk[j] & 0xffffu];
        y0[j] = L.y[(pk[j] >> 16) & 0xffffu];
        y1[j] = L.y[(pk[j] >> 32) & 0xffffu];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        const int t = t0 + j * NT;
        if (t < nt)
        {
          double d = kk[j];
          if ((pk[j] >> 48) == 0xffffu)
          {
            for (int i = m.term_dep_offsets[t]; i < m.term_dep_offsets[t + 1]; ++i)
              d *= L.y[m.term_dep_ids[i]];
          }
          else
          {
            d = packed_product(d, pk[j], y0[j], y1[j], 0.0);
          }
          L.jterm[t] = d;
        }
      }
    }
    coop_sync<NT>();
    for (int e = tid; e < m.nnz_jac; e += NT)
    {
      const double v = gather_sum(0.0, c.jel_term, c.jel_coef, c.jel_offsets[e], c.jel_offsets[e + 1], L.jterm);
      if constexpr (TO_ALU)
        L.alu[c.jac_to_alu[e]] = v;
      else
        L.jac[e] = v;
    }
    coop_sync<NT>();
  }

  // v - sum_q A[lo(terms[q])] * X[hi(terms[q])] for q in [b, e), subtracting in order;
  // with TC cells interleaved in A and X, position p of cell cc is p * TC + cc.
  // Terms go in batches of eight through a two-deep pipeline: while the subtractions
  // of batch i run in order, the operands of batch i + 1 and the table entries of
  // batch i + 2 are in flight.  Batches past the task's end read padding or the next
  // task's entries (the tables end in sixteen padding entries) and are not
  // subtracted; a short last batch skips the missing terms.
  template<int TC = 1>
  __device__ __forceinline__ double coop_dot(double v, const uint32_t* terms, uint32_t b, uint32_t e, const double* A,
                                             const double* X, int cc = 0)
  {
    if (b >= e)
      return v;
    uint32_t t[8];
    double a[8], x[8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
      t[r] = terms[b + r];
#pragma unroll
    for (int r = 0; r < 8; ++r)
    {
      a[r] = A[(t[r] & 0xffffu) * TC + cc];
      x[r] = X[(t[r] >> 16) * TC + cc];
    }
#pragma unroll
    for (int r = 0; r < 8; ++r)
      t[r] = terms[b + 8 + r];
    for (uint32_t q = b; q < e; q += 8)
    {
      double an[8], xn[8];
#pragma unroll
      for (int r = 0; r < 8; ++r)
      {
        an[r] = A[(t[r] & 0xffffu) * TC + cc];
        xn[r] = X[(t[r] >> 16) * TC + cc];
      }
      const uint32_t q2 = q + 16 < e ? q + 16 : q;
#pragma unroll
      for (int r = 0; r < 8; ++r)
        t[r] = terms[q2 + r];
      if (q + 8 <= e)
      {
#pragma unroll
        for (int r = 0; r < 8; ++r)
          v -= a[r] * x[r];
      }
      else
      {
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (q + r < e)
            v -= a[r] * x[r];
      }
#pragma unroll
      for (int r = 0; r < 8; ++r)
      {
        a[r] = an[r];
        x[r] = xn[r];
      }
    }
    return v;
  }

  // coop_dot for one task; a task of up to S terms reads its term entries, then all its
  // operands, at once
  template<int TC = 1, int S = 4>
  __device__ __forceinline__ double task_dot(double v, const uint32_t* terms, uint32_t b, uint32_t e, const double* A,
                                             const double* X, int cc = 0)
  {
    if (e - b > static_cast<uint32_t>(S))
      return coop_dot<TC>(v, terms, b, e, A, X, cc);
    uint32_t ts[S];
    double a[S], x[S];
#pragma unroll
    for (int r = 0; r < S; ++r)
      ts[r] = terms[b + r];
#pragma unroll
    for (int r = 0; r < S; ++r)
    {
      a[r] = A[(ts[r] & 0xffffu) * TC + cc];
      x[r] = X[(ts[r] >> 16) * TC + cc];
    }
#pragma unroll
    for (int r = 0; r < S; ++r)
      if (static_cast<uint32_t>(r) < e - b)
        v -= a[r] * x[r];
    return v;
  }

  // Workgroup barrier that waits for this wave's LDS operations only: global loads
  // issued before it stay in flight across it.
  template<int NT>
  __device__ __forceinline__ void lds_barrier()
  {
    if constexpr (NT == kWave)
    {
      coop_sync<kWave>();
    }
    else
    {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_s_waitcnt(0xc07f);  // vmcnt(63) expcnt(7) lgkmcnt(0)
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }

  // One-phase executor for schedules whose sums are split into short segments: each
  // thread runs whole tasks, a task of up to four terms reading all its operands at
  // once (longer tasks go through coop_dot).  The LU schedule is segmented into runs
  // of at least four products and the forward substitution's segments are short, so
  // this beats the two-phase form there; the back substitution keeps long sums and
  // runs two-phase.
  //
  // With GT (tables in global memory) the reads of a thread's first task are
  // pipelined over levels: while level l reads its operands, the same round of reads
  // fetches the term entries of the thread's task at level l + 1 and the task record
  // at level l + 2 (both constant), and the level barrier waits for LDS operations
  // only, so the table reads may take up to two levels to arrive.
  template<int NT, bool GT = false>
  __device__ inline void coop_levels_direct(const DevLevels& lv, const uint2* tasks, const uint32_t* terms,
                                            const double* A, const double* X, double* out)
  {
    const int tid = threadIdx.x;
    const int nl = lv.n_levels;
    if constexpr (!GT)
    {
      // tables in LDS: reading them level by level is as fast (measured) and keeps
      // fewer registers live
      for (int l = 0; l < nl; ++l)
      {
        const int tb = lv.offsets[l], te = lv.offsets[l + 1];
        for (int t = tb + tid; t < te; t += NT)
        {
          const uint2 task = tasks[t];
          const uint32_t e = tasks[t + 1].y;
          const uint32_t o = task.x & 0xffffu, d = task.x >> 16;
          const double dv = d < kUnitDivisor ? A[d] : 1.0;
          double v = task_dot(out[o], terms, task.y, e, A, X);
          if (d < kUnitDivisor)
            v = v / dv;
          out[o] = v;
        }
        coop_sync<NT>();
      }
      return;
    }
    if (nl == 0)
      return;
    auto bound = [&](int i) { return lv.offsets[i < nl ? i : nl]; };
    int b0 = bound(0), b1 = bound(1), b2 = bound(2), b3 = bound(3);
    // the thread's first task of the current level (record and terms) and of the next
    // level (record)
    bool c_has = b0 + tid < b1, n_has = b1 + tid < b2;
    uint2 c_task = make_uint2(0u, 0u), n_task = make_uint2(0u, 0u);
    uint32_t c_e = 0, n_e = 0, c_ts[4] = {0u, 0u, 0u, 0u};
    if (c_has)
    {
      c_task = tasks[b0 + tid];
      c_e = tasks[b0 + tid + 1].y;
    }
    if (n_has)
    {
      n_task = tasks[b1 + tid];
      n_e = tasks[b1 + tid + 1].y;
    }
    if (c_has)
    {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        c_ts[r] = terms[c_task.y + r];
    }
    for (int l = 0; l < nl; ++l)
    {
      const int b4 = bound(l + 4);
      // constant reads for the next two levels, issued with this level's operand reads
      const bool nn_has = b2 + tid < b3;
      uint2 nn_task = make_uint2(0u, 0u);
      uint32_t nn_e = 0, n_ts[4] = {0u, 0u, 0u, 0u};
      if (nn_has)
      {
        nn_task = tasks[b2 + tid];
        nn_e = tasks[b2 + tid + 1].y;
      }
      if (n_has)
      {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          n_ts[r] = terms[n_task.y + r];
      }
      for (int t = b0 + tid; t < b1; t += NT)
      {
        uint2 task = c_task;
        uint32_t e = c_e, ts[4] = {c_ts[0], c_ts[1], c_ts[2], c_ts[3]};
        if (t != b0 + tid)
        {
          task = tasks[t];
          e = tasks[t + 1].y;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            ts[r] = terms[task.y + r];
        }
        const uint32_t o = task.x & 0xffffu, d = task.x >> 16;
        const uint32_t n = e - task.y;
        double v = out[o];
        const double dv = d < kUnitDivisor ? A[d] : 1.0;
        if (n <= 4u)
        {
          double a[4], x[4];
#pragma unroll
          for (int r = 0; r < 4; ++r)
          {
            a[r] = A[ts[r] & 0xffffu];
            x[r] = X[ts[r] >> 16];
          }
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (static_cast<uint32_t>(r) < n)
              v -= a[r] * x[r];
        }
        else
        {
          v = coop_dot(v, terms, task.y, e, A, X);
        }
        if (d < kUnitDivisor)
          v = v / dv;
        out[o] = v;
      }
      c_has = n_has;
      c_task = n_task;
      c_e = n_e;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        c_ts[r] = n_ts[r];
      n_has = nn_has;
      n_task = nn_task;
      n_e = nn_e;
      b0 = b1;
      b1 = b2;
      b2 = b3;
      b3 = b4;
      if constexpr (GT)
        lds_barrier<NT>();
      else
        coop_sync<NT>();
    }
  }

  // A table that no thread writes during the kernel, read through the constant address
  // space: a wave-uniform index then becomes a scalar load (waited for with the LDS
  // counter at the next barrier, not with the vector-memory counter that the prefetched
  // records use).
  template<class T>
  __device__ __forceinline__ T const_load(const T* p, int i)
  {
    using CT = const __attribute__((address_space(4))) T;
    return ((CT*)(p))[i];
  }

  // A task's record (LevelSchedule::records): header and first kInlineTerms terms.
  struct TaskRecord
  {
    uint4 h, t1, t2, t3;  // h: out | diag << 16, first term, count, term 0; t1..t3: terms 1..12
  };
  static_assert(kRecordWords == 16, "a record is four uint4");

  // Makes every lane wait here for the reads that fill r, whether or not it then uses r:
  // a read left pending on a path that skips its use would make the wait before the next
  // write of r's registers (a whole level later) wait for every read issued since.
  __device__ __forceinline__ void settle(const uint4& r)
  {
    asm volatile("" ::"v"(r.x), "v"(r.y), "v"(r.z), "v"(r.w));
  }

  __device__ __forceinline__ TaskRecord load_record(const uint4* rec, int t)
  {
    TaskRecord r;
    r.h = rec[4 * t];
    r.t1 = rec[4 * t + 1];
    r.t2 = rec[4 * t + 2];
    r.t3 = rec[4 * t + 3];
    return r;
  }

  // One task from its record: out[o] = (out[o] - its products, subtracted in order) /
  // A[diag].  The first kInlineTerms products come with the record (five, then eight
  // more if the task has them); a longer task reads the rest from the term table.
  // PADDED: the record's slots past the task's last term name a zero in both A and X
  // (the LU's records), so every inline slot subtracts without a test (0 * 0 changes
  // nothing); otherwise each slot is tested, and slots 1 to 4 are skipped as a block
  // by a wavefront whose tasks have one term at most.
  template<bool PADDED>
  __device__ __forceinline__ void run_record(const TaskRecord& r, const uint32_t* terms, const double* A,
                                             const double* X, double* out)
  {
    const uint32_t o = r.h.x & 0xffffu, d = r.h.x >> 16, n = r.h.z;
    double v = out[o];
    const double dv = d < kUnitDivisor ? A[d] : 1.0;
    if constexpr (PADDED)
    {
      const uint32_t ts[5] = { r.h.w, r.t1.x, r.t1.y, r.t1.z, r.t1.w };
      double a[5], x[5];
#pragma unroll
      for (int q = 0; q < 5; ++q)
      {
        a[q] = A[ts[q] & 0xffffu];
        x[q] = X[ts[q] >> 16];
      }
#pragma unroll
      for (int q = 0; q < 5; ++q)
        v -= a[q] * x[q];
    }
    else
    {
      {
        const double a = A[r.h.w & 0xffffu], x = X[r.h.w >> 16];
        if (n > 0u)
          v -= a * x;
      }
      if (n > 1u)
      {
        const uint32_t ts[4] = { r.t1.x, r.t1.y, r.t1.z, r.t1.w };
        double a[4], x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
          a[q] = A[ts[q] & 0xffffu];
          x[q] = X[ts[q] >> 16];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (static_cast<uint32_t>(q) + 1u < n)
            v -= a[q] * x[q];
      }
    }
    if (n > 5u)
    {
      const uint32_t ts[8] = { r.t2.x, r.t2.y, r.t2.z, r.t2.w, r.t3.x, r.t3.y, r.t3.z, r.t3.w };
      double a[8], x[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
      {
        a[q] = A[ts[q] & 0xffffu];
        x[q] = X[ts[q] >> 16];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (PADDED || static_cast<uint32_t>(q) + 5u < n)
          v -= a[q] * x[q];
    }
    // the rest in batches of eight, each batch's reads consumed before the next batch's
    // (reads past the task's last term stay inside the term table's padding), so that
    // no read is left in flight when the task ends
    for (uint32_t q = r.h.y + kInlineTerms, e = r.h.y + n; q < e; q += 8)
    {
      uint32_t ts[8];
      double a[8], x[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        ts[k] = terms[q + k];
#pragma unroll
      for (int k = 0; k < 8; ++k)
      {
        a[k] = A[ts[k] & 0xffffu];
        x[k] = X[ts[k] >> 16];
      }
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (q + k < e)
          v -= a[k] * x[k];
    }
    if (d < kUnitDivisor)
      v = v / dv;
    out[o] = v;
  }

  __device__ __forceinline__ double readlane_f64(double v, uint32_t lane)
  {
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b), lane));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane));
    return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
  }

  // v - p[0] - p[1] - ... - p[n-1], subtracted in that order, in batches of eight whose
  // reads are issued before the previous batch is subtracted; reads up to fifteen
  // entries past p[n-1] (unused), which must lie in LDS the caller owns.
  __device__ __forceinline__ double ordered_sub(double v, const double* p, uint32_t n)
  {
    double a[8], b[8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
      a[r] = p[r];
    uint32_t q = 0;
    for (; q + 16 <= n; q += 16)
    {
#pragma unroll
      for (int r = 0; r < 8; ++r)
        b[r] = p[q + 8 + r];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < 8; ++r)
        v -= a[r];
#pragma unroll
      for (int r = 0; r < 8; ++r)
        a[r] = p[q + 16 + r];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < 8; ++r)
        v -= b[r];
    }
    if (q + 8 <= n)
    {
#pragma unroll
      for (int r = 0; r < 8; ++r)
        b[r] = p[q + 8 + r];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < 8; ++r)
        v -= a[r];
#pragma unroll
      for (int r = 0; r < 8; ++r)
        a[r] = b[r];
      q += 8;
    }
#pragma unroll
    for (int r = 0; r < 8; ++r)
      if (q + r < n)
        v -= a[r];
    return v;
  }

  // v - p[0] - ... - p[n-1] in that order for n <= 64, straight-line code: batch k + 1's
  // reads are issued before batch k is subtracted, and the batch that holds the last
  // products subtracts only those (reads up to fifteen entries past p[n-1], unused, in
  // LDS the caller owns).
  __device__ __forceinline__ double ordered_sub_short(double v, const double* p, uint32_t n)
  {
    const uint32_t nb = n >> 3, rem = n & 7u;
    double cur[8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
      cur[r] = p[r];
#pragma unroll
    for (uint32_t k = 0; k <= 8; ++k)
    {
      if (k == nb)
      {
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r)
          if (r < rem)
            v -= cur[r];
        break;
      }
      double nxt[8];
#pragma unroll
      for (int r = 0; r < 8; ++r)
        nxt[r] = p[8 * k + 8 + r];
#pragma unroll
      for (int r = 0; r < 8; ++r)
        v -= cur[r];
#pragma unroll
      for (int r = 0; r < 8; ++r)
        cur[r] = nxt[r];
    }
    return v;
  }

  // What the back substitution's task chain needs before it starts, read at the start of
  // the solve: this lane's header words of tasks `lane` and lane + 64, and its term words
  // of tasks 0 to 4.
  struct ChainAhead
  {
    uint2 h0, h1;
    uint32_t t0;
    uint32_t terms[4];
  };

  __device__ __forceinline__ ChainAhead chain_ahead(const DevChain& ch, bool on)
  {
    ChainAhead a{};
    if (on && threadIdx.x < kWave)
    {
      const uint32_t lane = threadIdx.x;
      const uint2* h = reinterpret_cast<const uint2*>(ch.headers);
      a.h0 = h[lane];
      a.h1 = h[kWave + lane];
      const int last = ch.n_tasks - 1;
      a.t0 = ch.lane_terms[lane];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        a.terms[q] = ch.lane_terms[static_cast<size_t>(1 + q < last ? 1 + q : last) * kChainLanes + lane];
    }
    return a;
  }

  // The back substitution's task chain (TaskChain) by the first wavefront, tasks in
  // order.  Lane k reads the operands of product k of task t + 1 before task t is summed
  // (all lanes alike, from the products buffer of its parity) and writes the product to
  // the other buffer after it, except the late one (reading task t's result), whose
  // factor value it keeps; the sum of task t + 1 forms that product from the result in a
  // register.  Tasks run in blocks of four; the lane term words of the next block are
  // read when a block starts.  P holds two buffers of kChainLanes doubles and the
  // ordered sums' read slack.
  template<int NT>
  __device__ inline void coop_chain(const DevChain& ch, const ChainAhead& ahead, const double* A, double* x,
                                    double* P)
  {
    if (threadIdx.x < kWave)
    {
      const uint32_t lane = threadIdx.x;
      const int nt = ch.n_tasks;
      auto term_of = [&](int t) { return ch.lane_terms[static_cast<size_t>(t < nt ? t : nt - 1) * kChainLanes + lane]; };
      // header words of task t (uniform)
      auto head = [&](int t, bool second) {
        const uint32_t w0 = second ? ahead.h0.y : ahead.h0.x, w1 = second ? ahead.h1.y : ahead.h1.x;
        return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(t < kWave ? w0 : w1), t & (kWave - 1)));
      };
      // products of task 0 (nothing before it in the chain)
      {
        const double a = A[ahead.t0 & 0xffffu], xx = x[ahead.t0 >> 16];
        if (lane < (head(0, true) & 0xffffu))
          P[lane] = a * xx;
      }
      double x_prev = 0.0, late_now = 0.0;
      auto step = [&](int t, uint32_t term_next) {
        if (t >= nt)
          return;
        // the operands of task t + 1's products (none past the last task): read first,
        // multiplied after task t, so that their reads overlap its sum
        const uint32_t w1 = t + 1 < nt ? head(t + 1, true) : 0u;
        const uint32_t n1 = w1 & 0xffffu, late1 = w1 >> 16;
        const double a1 = A[term_next & 0xffffu], x1 = x[term_next >> 16];
        // task t
        const uint32_t od = head(t, false), w = head(t, true);
        const uint32_t o = od & 0xffffu, d = od >> 16, n = w & 0xffffu, late = w >> 16;
        const double* p = P + (t & 1) * kChainLanes;
        const double dv = d < kUnitDivisor ? A[d] : 1.0;
        double v = x[o];
        __builtin_amdgcn_sched_barrier(0);
        if (late < n)
        {
          if (late > 0)
            v = ordered_sub_short(v, p, late);
          v -= readlane_f64(late_now, late) * x_prev;
          v = ordered_sub_short(v, p + late + 1, n - late - 1);
        }
        else
        {
          v = ordered_sub_short(v, p, n);
        }
        if (d < kUnitDivisor)
          v = v / dv;
        if (lane == 0)
          x[o] = v;
        x_prev = v;
        __builtin_amdgcn_sched_barrier(0);
        // task t + 1's products that do not read task t's result; the late one's factor
        late_now = a1;
        if (lane < n1 && lane != late1)
          P[((t + 1) & 1) * kChainLanes + lane] = a1 * x1;
      };
      uint32_t cur[4] = { ahead.terms[0], ahead.terms[1], ahead.terms[2], ahead.terms[3] };
      for (int t = 0; t < nt; t += 4)
      {
        uint32_t nxt[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          nxt[q] = term_of(t + 5 + q);
        step(t, cur[0]);
        step(t + 1, cur[1]);
        step(t + 2, cur[2]);
        step(t + 3, cur[3]);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          cur[q] = nxt[q];
      }
    }
    lds_barrier<NT>();
  }

  // The forward substitution's column sweep (ColumnSweep) by the first wavefront: lane r
  // keeps its row's partial sum in a register; at column c the column's value comes from
  // the lane that finished it (or from x, final before the sweep) through a lane read,
  // and every lane with a term there subtracts L * x[c] (the others keep their sum by a
  // select, so no value of x enters a row it is not in).  A lane's swept terms sit at
  // consecutive ALU positions, its term at column c at start + (its columns below c);
  // each factor value is read two columns before its use, over three rotating registers
  // (unrolled by three so that none moves), the reads for columns a lane has no term at
  // staying inside the factor and its zero.  `me` is this lane's entry of the table,
  // read before the levels that precede the sweep.
  template<int NT>
  __device__ inline void coop_sweep(const DevSweep& sw, const uint4& me, const double* A, double* x)
  {
    if (threadIdx.x < kWave)
    {
      const uint32_t lane = threadIdx.x;
      const int n_cols = sw.n_cols;
      const uint32_t row = me.x & 0xffffu, start = me.x >> 16;
      const uint64_t mask = me.y | (static_cast<uint64_t>(me.z) << 32);
      const uint32_t src = me.w;
      double v = row != 0xffffu ? x[row] : 0.0;
      const double xv = lane < static_cast<uint32_t>(n_cols) && !(src & kSweepFromLane) ? x[src] : 0.0;
      auto factor = [&](int c) {
        const int cc = c < 63 ? c : 63;
        return A[start + static_cast<uint32_t>(__popcll(mask & ((uint64_t(1) << cc) - 1u)))];
      };
      auto step = [&](int c, double a) {
        const uint32_t cl = static_cast<uint32_t>(c) & 63u;
        const uint32_t s = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(src), cl));
        // both lane reads, then a scalar select (no branch per column)
        const double from_lane = readlane_f64(v, s & 63u), from_x = readlane_f64(xv, cl);
        const double xc = (s & kSweepFromLane) ? from_lane : from_x;
        const double nv = v - a * xc;
        v = (c < n_cols && ((mask >> cl) & 1u)) ? nv : v;
      };
      double a0 = factor(0), a1 = factor(1), a2;
      for (int c = 0; c < n_cols; c += 3)
      {
        a2 = factor(c + 2);
        step(c, a0);
        a0 = factor(c + 3);
        step(c + 1, a1);
        a1 = factor(c + 4);
        step(c + 2, a2);
      }
      if (row != 0xffffu)
        x[row] = v;
    }
    lds_barrier<NT>();
  }

  // A level schedule with its records in global memory, pipelined: at the start of level
  // l each thread reads the record of its first task at level l + 2, then runs level l
  // from the record read two levels earlier, so every read has a whole level to arrive.
  // Three buffers rotate over levels l, l + 1 and l + 2, unrolled by three so that each
  // keeps its registers (a rotation by register moves would wait for the read just
  // issued).  A thread without a task at a level reads the sentinel record.  Level
  // bounds are scalar reads (constant address space).  Same operations in the same order
  // as coop_levels_direct.
  template<int NT, bool PADDED>
  __device__ inline void coop_levels_piped(const DevLevels& lv, const uint32_t* terms, const double* A,
                                           const double* X, double* out)
  {
    const int tid = threadIdx.x;
    const int nl = lv.n_levels;
    if (nl == 0)
      return;
    const uint4* rec = lv.records;
    auto bound = [&](int i) { return const_load(lv.offsets, i < nl ? i : nl); };
    auto first_record = [&](int l, int b, int e) {
      return load_record(rec, (l < nl && b + tid < e) ? b + tid : lv.n_tasks);
    };
    int b0 = bound(0), b1 = bound(1), b2 = bound(2), b3 = bound(3);
    TaskRecord r0 = first_record(0, b0, b1), r1 = first_record(1, b1, b2), r2;
    auto step = [&](int l, const TaskRecord& cur, TaskRecord& ahead) {
      const int b4 = bound(l + 4);
      ahead = first_record(l + 2, b2, b3);
      settle(cur.h);
      settle(cur.t1);
      settle(cur.t2);
      settle(cur.t3);
      if (b0 + tid < b1)
        run_record<PADDED>(cur, terms, A, X, out);
      for (int t = b0 + tid + NT; t < b1; t += NT)
        run_record<PADDED>(load_record(rec, t), terms, A, X, out);
      if (l < nl)
        lds_barrier<NT>();
      b0 = b1;
      b1 = b2;
      b2 = b3;
      b3 = b4;
    };
    // steps past the last level have no tasks and no barrier (a break between the steps
    // would join the paths in a way that makes the next wait cover every read)
    for (int l = 0; l < nl; l += 3)
    {
      step(l, r0, r2);
      step(l + 1, r1, r0);
      step(l + 2, r2, r1);
    }
  }

  // The two-phase form of coop_levels with its records and terms in global memory,
  // pipelined like coop_levels_piped: the first product term and the first task record
  // of level l + 2 are read at the start of level l.  Needs every level's products to
  // fit P.
  template<int NT>
  __device__ inline void coop_levels_piped2(const DevLevels& lv, const uint32_t* terms, const double* A,
                                            const double* X, double* out, double* P)
  {
    const int tid = threadIdx.x;
    const int nl = lv.n_levels;
    if (nl == 0)
      return;
    const uint4* rec = lv.records;
    auto bound = [&](int i) { return const_load(lv.offsets, i < nl ? i : nl); };
    // first term of level i (the end of the terms past the last level)
    auto term_bound = [&](int i) {
      return static_cast<int>(const_load(reinterpret_cast<const uint32_t*>(lv.tasks), 2 * bound(i) + 1));
    };
    struct Ahead
    {
      uint4 h;      // first task's record header (its terms are not needed: the sums read P)
      uint32_t pt;  // first product term
    };
    auto fetch = [&](int l, int b, int e, int qb, int qe) {
      Ahead a;
      a.h = rec[4 * ((l < nl && b + tid < e) ? b + tid : lv.n_tasks)];
      a.pt = terms[(l < nl && qb + tid < qe) ? qb + tid : 0];
      return a;
    };
    int b0 = bound(0), b1 = bound(1), b2 = bound(2), b3 = bound(3);
    int q0 = term_bound(0), q1 = term_bound(1), q2 = term_bound(2), q3 = term_bound(3);
    Ahead r0 = fetch(0, b0, b1, q0, q1), r1 = fetch(1, b1, b2, q1, q2), r2;
    // a task's ordered sum of its products
    auto sum = [&](const uint4& h) {
      const uint32_t o = h.x & 0xffffu, d = h.x >> 16, n = h.z;
      const double dv = d < kUnitDivisor ? A[d] : 1.0;
      double v = ordered_sub(out[o], P + (static_cast<int>(h.y) - q0), n);
      if (d < kUnitDivisor)
        v = v / dv;
      out[o] = v;
    };
    auto step = [&](int l, const Ahead& cur, Ahead& ahead) {
      const int b4 = bound(l + 4), q4 = term_bound(l + 4);
      ahead = fetch(l + 2, b2, b3, q2, q3);
      settle(cur.h);
      settle(uint4{ cur.pt, 0u, 0u, 0u });
      if (q0 + tid < q1)
        P[tid] = A[cur.pt & 0xffffu] * X[cur.pt >> 16];
      for (int q = q0 + tid + NT; q < q1; q += NT)
      {
        const uint32_t tr = terms[q];
        P[q - q0] = A[tr & 0xffffu] * X[tr >> 16];
      }
      if (l < nl)
        lds_barrier<NT>();
      if (b0 + tid < b1)
        sum(cur.h);
      for (int t = b0 + tid + NT; t < b1; t += NT)
        sum(rec[4 * t]);
      if (l < nl)
        lds_barrier<NT>();
      b0 = b1;
      b1 = b2;
      b2 = b3;
      b3 = b4;
      q0 = q1;
      q1 = q2;
      q2 = q3;
      q3 = q4;
    };
    // steps past the last level have no tasks and no barrier (a break between the steps
    // would join the paths in a way that makes the next wait cover every read)
    for (int l = 0; l < nl; l += 3)
    {
      step(l, r0, r2);
      step(l + 1, r1, r0);
      step(l + 2, r2, r1);
    }
  }

  // Runs a level schedule: task t computes out[o] = (out[o] - dot) / A[diag] (no
  // division when diag is 0xffff), reading A and X.
  //
  // A level whose products fit the buffer P (p_capacity doubles) runs in two phases:
  // every thread forms products of the level (the level's terms are contiguous), then
  // one thread per task subtracts its products from P in order.  Each product is
  // rounded once either way, so the result is that of subtracting a * x term by
  // term; the long sequential part reads a contiguous stream instead of two indexed
  // operands per term.  Larger levels run task by task.
  //
  // V selects code shapes that give identical results (bit 0: products eight per
  // thread per round on levels of more than 2 NT terms -- the small levels of the
  // substitutions run faster one product per round; bit 1: each thread's first task
  // loaded before the products; bit 2: sums in batches of sixteen, else eight;
  // bit 3: the tail of a sum in batches of four).
  constexpr int kLevelShape = 1;

  template<int NT, int V = kLevelShape>
  __device__ inline void coop_levels(const DevLevels& lv, const uint2* tasks, const uint32_t* terms, const double* A,
                                     const double* X, double* out, double* P, int p_capacity)
  {
    const int tid = threadIdx.x;
    for (int l = 0; l < lv.n_levels; ++l)
    {
      const int tb = lv.offsets[l], te = lv.offsets[l + 1];
      const uint32_t qb = tasks[tb].y, qe = tasks[te].y;
      if (static_cast<int>(qe - qb) <= p_capacity)
      {
        const int t0 = tb + tid;
        uint2 task0 = make_uint2(0u, 0u);
        uint32_t e0 = 0;
        double v0 = 0.0, d0 = 1.0;
        if constexpr ((V & 2) != 0)
        {
          if (t0 < te)
          {
            task0 = tasks[t0];
            e0 = tasks[t0 + 1].y;
            v0 = out[task0.x & 0xffffu];
            if ((task0.x >> 16) < kUnitDivisor)
              d0 = A[task0.x >> 16];
          }
        }
        if ((V & 1) != 0 && qe - qb > 2u * NT)
        {
          for (uint32_t q0 = qb + tid; q0 < qe; q0 += 8 * NT)
          {
            uint32_t tr[8];
            double a[8], x[8];
#pragma unroll
            for (int r = 0; r < 8; ++r)
            {
              const uint32_t q = q0 + r * NT;
              tr[r] = terms[q < qe ? q : qb];
            }
#pragma unroll
            for (int r = 0; r < 8; ++r)
            {
              a[r] = A[tr[r] & 0xffffu];
              x[r] = X[tr[r] >> 16];
            }
#pragma unroll
            for (int r = 0; r < 8; ++r)
            {
              const uint32_t q = q0 + r * NT;
              if (q < qe)
                P[q - qb] = a[r] * x[r];
            }
          }
        }
        else
        {
          for (uint32_t q = qb + tid; q < qe; q += NT)
          {
            const uint32_t tr = terms[q];
            P[q - qb] = A[tr & 0xffffu] * X[tr >> 16];
          }
        }
        coop_sync<NT>();
        for (int t = t0; t < te; t += NT)
        {
          uint2 task;
          uint32_t e;
          double v, dv = 1.0;
          if ((V & 2) != 0 && t == t0)
          {
            task = task0;
            e = e0;
            v = v0;
            dv = d0;
          }
          else
          {
            task = tasks[t];
            e = tasks[t + 1].y;
            v = out[task.x & 0xffffu];
            if ((task.x >> 16) < kUnitDivisor)
              dv = A[task.x >> 16];
          }
          const double* p = P + (task.y - qb);
          const uint32_t n = e - task.y;
          uint32_t q = 0;
          constexpr int kB = (V & 4) != 0 ? 16 : 8;
          for (; q + kB <= n; q += kB)
          {
            double s[kB];
#pragma unroll
            for (int r = 0; r < kB; ++r)
              s[r] = p[q + r];
#pragma unroll
            for (int r = 0; r < kB; ++r)
              v -= s[r];
          }
          if constexpr ((V & 8) != 0)
          {
            for (; q + 4 <= n; q += 4)
            {
              double s[4];
#pragma unroll
              for (int r = 0; r < 4; ++r)
                s[r] = p[q + r];
#pragma unroll
              for (int r = 0; r < 4; ++r)
                v -= s[r];
            }
          }
          for (; q < n; ++q)
            v -= p[q];
          if ((task.x >> 16) < kUnitDivisor)
            v = v / dv;
          out[task.x & 0xffffu] = v;
        }
      }
      else
      {
        for (int t = tb + tid; t < te; t += NT)
        {
          const uint2 task = tasks[t];
          const uint32_t qe2 = tasks[t + 1].y;
          const uint32_t o = task.x & 0xffffu, d = task.x >> 16;
          double v = coop_dot(out[o], terms, task.y, qe2, A, X);
          if (d < kUnitDivisor)
            v = v / A[d];
          out[o] = v;
        }
      }
      coop_sync<NT>();
    }
  }

  // Table pointers for one launch: global memory or the workgroup's LDS copy.
  struct CoopTables
  {
    const uint2 *lu_tasks, *fwd_tasks, *bwd_tasks;
    const uint32_t *lu_terms, *fwd_terms, *bwd_terms;
  };

  // In-place LU of L.jac into L.alu.
  template<int NT, bool GT = false>
  __device__ inline void coop_decompose(const DevMech& m, const DevCoop& c, const CoopLds& L, const CoopTables& T)
  {
    const int tid = threadIdx.x;
    // the factor and the zero one past it (that the LU records' padded slots read)
    for (int q = tid; q <= c.nnz_alu; q += NT)
      L.alu[q] = 0.0;
    coop_sync<NT>();
    for (int e = tid; e < m.nnz_jac; e += NT)
      L.alu[c.jac_to_alu[e]] = L.jac[e];
    coop_sync<NT>();
    if constexpr (GT)
      coop_levels_piped<NT, true>(c.lu, c.lu.terms, L.alu, L.alu, L.alu);
    else
      coop_levels_direct<NT, false>(c.lu, T.lu_tasks, T.lu_terms, L.alu, L.alu, L.alu);
  }

  // x <- U^-1 L^-1 x with the factors in L.alu (unit lower diagonal).
  template<int NT, bool GT = false>
  __device__ inline void coop_solve(const DevCoop& c, const CoopLds& L, const CoopTables& T, double* x)
  {
    // the forward substitution's levels before its column sweep, then the sweep
    const bool sweep = c.fwd_sweep.first_level >= 0;
    const uint4 me = sweep && threadIdx.x < kWave ? c.fwd_sweep.lanes[threadIdx.x] : make_uint4(0u, 0u, 0u, 0u);
    // what the back substitution's task chain reads first
    const bool chain = GT && c.bwd_chain.levels > 0 && L.p_capacity >= 2 * kChainLanes;
    const ChainAhead ahead = chain_ahead(c.bwd_chain, chain);
    DevLevels fwd = c.fwd;
    if (sweep)
      fwd.n_levels = c.fwd_sweep.first_level;
    if constexpr (GT)
    {
      coop_levels_piped<NT, false>(fwd, c.fwd.terms, L.alu, x, x);
      if (sweep)
        coop_sweep<NT>(c.fwd_sweep, me, L.alu, x);
      // the back substitution's first levels as a task chain, then the rest by levels
      DevLevels bwd = c.bwd;
      if (chain)
      {
        coop_chain<NT>(c.bwd_chain, ahead, L.alu, x, L.products);
        bwd.offsets += c.bwd_chain.levels;
        bwd.n_levels -= c.bwd_chain.levels;
      }
      if (c.bwd.max_level_terms <= L.p_capacity)
        coop_levels_piped2<NT>(bwd, c.bwd.terms, L.alu, x, x, L.products);
      else
        coop_levels<NT>(bwd, T.bwd_tasks, T.bwd_terms, L.alu, x, x, L.products, L.p_capacity);
    }
    else
    {
      coop_levels_direct<NT, false>(fwd, T.fwd_tasks, T.fwd_terms, L.alu, x, x);
      if (sweep)
        coop_sweep<NT>(c.fwd_sweep, me, L.alu, x);
      coop_levels<NT>(c.bwd, T.bwd_tasks, T.bwd_terms, L.alu, x, x, L.products, L.p_capacity);
    }
  }

  // 32-bit words of the tables a launch stages in LDS.
  __host__ __device__ inline size_t coop_solve_table_words(const DevCoop& c)
  {
    return 2 * (static_cast<size_t>(c.fwd.n_tasks) + 1) + c.fwd.n_terms + 2 * (static_cast<size_t>(c.bwd.n_tasks) + 1) +
           c.bwd.n_terms;
  }

  __host__ __device__ inline size_t coop_lu_table_words(const DevCoop& c)
  {
    return 2 * (static_cast<size_t>(c.lu.n_tasks) + 1) + c.lu.n_terms;
  }
}  // namespace micm_hip
