// solver.hip -- HIP kernels of the Rosenbrock path and the C-ABI of include/micm_hip.h.
//
// Layout: one lane = one grid cell; per-cell arrays in the grouped layout
// ((cell / 64) * n_values + v) * 64 + cell % 64.  Kernels launch 256-thread
// workgroups (4 wavefronts = 4 cell groups).
//
// Integration forms (micm_hip_solve):
//   LOCKSTEP   -- every cell of the state shares one step size, as the reference's
//                 multi-cell State (rosenbrock.inl:12-304); the normalized error is
//                 the reference's single running sum over cells and species.  A small
//                 State runs the whole loop in one workgroup (k_lockstep_fused).
//   PLAIN      -- per-cell step size; one launch per phase of an attempt over all
//                 cells (tiled gathers, factorisation and substitutions; per-cell or
//                 per-(cell, species) kernels for the rest), host loop.
//   PERSISTENT -- per-cell step size; one persistent launch whose wavefronts take
//                 cells from a work queue and carry each to the end of its step.
//   COOPERATIVE -- per-cell step size; one workgroup per cell (coop.hpp), cells from
//                 a work queue, in the previous Solve's cost order once there was one.
#include "adjoint.hpp"
#include "coop.hpp"
#include "device.hpp"
#include "mechanism.hpp"
#include "path.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

using namespace micm_hip;

namespace
{
  thread_local std::string g_last_error = "";

  int fail(int code, const std::string& msg)
  {
    g_last_error = msg;
    return code;
  }

#define HIP_TRY(expr)                                                                                     \
  do                                                                                                      \
  {                                                                                                       \
    hipError_t _e = (expr);                                                                               \
    if (_e != hipSuccess)                                                                                 \
      return fail(MICM_HIP_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));                \
  } while (0)

  constexpr int kBlock = 256;

  __host__ __device__ inline int64_t n_groups_of(int64_t n_cells)
  {
    return (n_cells + kGroup - 1) / kGroup;
  }

  inline unsigned grid_for(int64_t n_cells)
  {
    return static_cast<unsigned>((n_groups_of(n_cells) * kGroup + kBlock - 1) / kBlock);
  }

  // reaction chunks of the rate-constant kernel: at least 32 reactions each, and enough
  // chunks for about 8192 waves in all, at most 16
  inline dim3 rate_grid(int64_t n_cells, int n_reactions)
  {
    const int64_t waves = n_groups_of(n_cells);
    int chunks = static_cast<int>(std::min<int64_t>(16, std::max<int64_t>(1, (8192 + waves - 1) / waves)));
    chunks = std::max(1, std::min(chunks, (n_reactions + 31) / 32));
    return dim3(grid_for(n_cells), static_cast<unsigned>(chunks));
  }

  // ------------------------------------------------------------------ kernel-level entry points

  // one cell per lane; the reactions are split into gridDim.y contiguous chunks so that a
  // launch over few cells still has enough waves to hide the transcendental latency
  __global__ void __launch_bounds__(kBlock) k_rate_constants(DevMech m, int64_t n_cells, const double* __restrict__ cond,
                                                             const double* __restrict__ custom, double* __restrict__ k)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    const size_t g = c / kGroup;
    const int l = c % kGroup;
    Lane cd = lane_view(const_cast<double*>(cond), g, 3, l);
    Lane cu = lane_view(const_cast<double*>(custom), g, m.n_custom, l);
    Lane kk = lane_view(k, g, m.n_reactions, l);
    const double T = cd[0], P = cd[1], rho = cd[2];
    const int per = (m.n_reactions + gridDim.y - 1) / gridDim.y;
    const int rb = blockIdx.y * per, re = min(m.n_reactions, rb + per);
    for (int r = rb; r < re; ++r)
      kk[r] = rate_constant(m, r, T, P, rho, cu);
  }

  // FROM_ZERO: the output is cleared first (the reference's Fill(0) before the call)
  template<bool FROM_ZERO = false>
  __global__ void __launch_bounds__(kBlock) k_add_forcing(DevMech m, int64_t n_cells, const double* __restrict__ k,
                                                          const double* __restrict__ y, double* __restrict__ f)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    const size_t g = c / kGroup;
    const int l = c % kGroup;
    auto out = lane_view(f, g, m.n_species, l);
    if constexpr (FROM_ZERO)
      for (int i = 0; i < m.n_species; ++i)
        out[i] = 0.0;
    add_forcing(m, lane_view(const_cast<double*>(k), g, m.n_reactions, l), lane_view(const_cast<double*>(y), g, m.n_species, l),
                out);
  }

  template<bool FROM_ZERO = false>
  __global__ void __launch_bounds__(kBlock) k_subtract_jacobian(DevMech m, int64_t n_cells, const double* __restrict__ k,
                                                                const double* __restrict__ y, double* __restrict__ jac)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    const size_t g = c / kGroup;
    const int l = c % kGroup;
    auto out = lane_view(jac, g, m.nnz_jac, l);
    if constexpr (FROM_ZERO)
      for (int i = 0; i < m.nnz_jac; ++i)
        out[i] = 0.0;
    subtract_jacobian(m, lane_view(const_cast<double*>(k), g, m.n_reactions, l),
                      lane_view(const_cast<double*>(y), g, m.n_species, l), out);
  }

  // ---- tiled single-phase kernels: one workgroup per TC cells of a group, the tile's
  // y and k staged in LDS, outputs gathered per species / Jacobian element in the
  // reference's accumulation order and written once.  Each HBM byte of y, k and the
  // output moves once.
  //
  // Workgroups are dealt to the eight XCDs round-robin (workgroup b runs on XCD b % 8),
  // and each XCD has its own L2.  The tile of workgroup b is chosen so that the 64 / TC
  // tiles of one group run on one XCD one after another and share the group's
  // 128-byte lines in that L2.
  constexpr int kXcds = 8;
  constexpr int kTileThreads = 256;

  template<int TC>
  __host__ __device__ constexpr int tiles_per_group()
  {
    return kGroup / TC;
  }

  template<int TC>
  __device__ inline bool tile_of_block(int64_t n_cells, size_t& g, int& q0)
  {
    constexpr int S = tiles_per_group<TC>();
    const unsigned b = blockIdx.x, x = b % kXcds, j = b / kXcds;
    const int64_t grp = static_cast<int64_t>(j / S) * kXcds + x;
    g = static_cast<size_t>(grp);
    q0 = static_cast<int>(j % S) * TC;
    return grp < n_groups_of(n_cells);
  }

  // dst[v * TC + cc] = src[(g * n_values + v) * 64 + q0 + cc], eight loads in flight per thread
  template<int TC, int NT>
  __device__ inline void load_tile(double* dst, const double* __restrict__ src, size_t g, int n_values, int q0)
  {
    const int total = n_values * TC;
    const double* base = src + g * static_cast<size_t>(n_values) * kGroup + q0;
    for (int i0 = threadIdx.x; i0 < total; i0 += 8 * NT)
    {
      double v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
      {
        const int i = i0 + j * NT < total ? i0 + j * NT : i0;
        v[j] = base[(i / TC) * kGroup + i % TC];
      }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (i0 + j * NT < total)
          dst[i0 + j * NT] = v[j];
    }
  }

  // rates of the tile's cells, in place over their rate constants, four reactions per
  // thread at a time with packed reactant ids
  template<int TC, int NT>
  __device__ inline void tile_rates(const DevMech& m, const DevCoop& c, const double* yt, double* rt)
  {
    const int nr = m.n_reactions * TC;
    const int cc = threadIdx.x % TC;  // NT is a multiple of TC
    for (int i0 = threadIdx.x; i0 < nr; i0 += 4 * NT)
    {
      uint64_t pk[4];
      double kk[4], y0[4], y1[4], y2[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        const int i = i0 + j * NT < nr ? i0 + j * NT : i0;
        pk[j] = c.rxn_pack[i / TC];
        kk[j] = rt[i];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        y0[j] = yt[(pk[j] & 0xffffu) * TC + cc];
        y1[j] = yt[((pk[j] >> 16) & 0xffffu) * TC + cc];
        y2[j] = yt[((pk[j] >> 32) & 0xffffu) * TC + cc];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        const int i = i0 + j * NT;
        if (i < nr)
        {
          double rate = kk[j];
          if ((pk[j] >> 48) == 0xffffu)
          {
            const int r = i / TC;
            for (int a = m.reactant_offsets[r]; a < m.reactant_offsets[r + 1]; ++a)
              rate *= yt[m.reactant_ids[a] * TC + cc];
          }
          else
          {
            rate = packed_product(rate, pk[j], y0[j], y1[j], y2[j]);
          }
          rt[i] = rate;
        }
      }
    }
  }

  // f += f(y): rates of the tile's cells into LDS (in place over k), then per-species
  // gathers; the old value of the next item is loaded while the current one is summed.
  // FROM_ZERO: f = f(y), the reference's Fill(0) and AddForcingTerms in one pass (f is only
  // written).
  template<int TC, int NT, bool FROM_ZERO = false>
  __global__ void __launch_bounds__(NT) k_add_forcing_tiled(DevMech m, DevCoop c, int64_t n_cells,
                                                            const double* __restrict__ k,
                                                            const double* __restrict__ y, double* __restrict__ f)
  {
    extern __shared__ double tile[];
    size_t g;
    int q0;
    if (!tile_of_block<TC>(n_cells, g, q0))
      return;
    const int N = m.n_species, R = m.n_reactions;
    double* yt = tile;
    double* rt = tile + static_cast<size_t>(N) * TC;
    load_tile<TC, NT>(yt, y, g, N, q0);
    load_tile<TC, NT>(rt, k, g, R, q0);
    __syncthreads();
    tile_rates<TC, NT>(m, c, yt, rt);
    __syncthreads();
    const int cc = threadIdx.x % TC;
    if (static_cast<int64_t>(g) * kGroup + q0 + cc >= n_cells)
      return;
    const int ns = N * TC;
    double* out = f + g * static_cast<size_t>(N) * kGroup + q0 + cc;
    int i = threadIdx.x;
    double v_next = !FROM_ZERO && i < ns ? out[(i / TC) * kGroup] : 0.0;
    for (; i < ns; i += NT)
    {
      const int sp = i / TC;
      double v = v_next;
      if (!FROM_ZERO && i + NT < ns)
        v_next = out[((i + NT) / TC) * kGroup];
      v = gather_sum<TC>(v, c.spec_rxn, c.spec_coef, c.spec_offsets[sp], c.spec_offsets[sp + 1], rt, cc);
      out[sp * kGroup] = v;
    }
  }

  // jac -= J(y): per Jacobian element, its terms' products k * y... in the reference's
  // order, two entries at a time with packed operands (elements average one to two
  // terms).  FROM_ZERO: jac = -J(y), the reference's Fill(0) and SubtractJacobianTerms in
  // one pass (jac is only written).  DUE: only the running cells whose control flags ask
  // for a new Jacobian, each formed from zero (the per-cell forms' step start); the other
  // cells' Jacobians are left as they are.
  template<int TC, int NT, bool DUE, bool FROM_ZERO = DUE>
  __device__ inline void subtract_jacobian_tile(const DevMech& m, const DevCoop& c, int64_t n_cells,
                                                const double* __restrict__ k, const double* __restrict__ y,
                                                double* __restrict__ jac, const CellControl* __restrict__ ctl)
  {
    extern __shared__ double tile[];
    size_t g;
    int q0;
    if (!tile_of_block<TC>(n_cells, g, q0))
      return;
    const int cc = threadIdx.x % TC;
    const int64_t cell = static_cast<int64_t>(g) * kGroup + q0 + cc;
    bool due = cell < n_cells;
    if constexpr (DUE)
    {
      due = due && ctl[cell].status == MICM_HIP_STATE_RUNNING && (ctl[cell].flags & kJacobianDue) != 0;
      if (!__syncthreads_or(due ? 1 : 0))
        return;
    }
    const int N = m.n_species, R = m.n_reactions;
    double* yt = tile;
    double* kt = tile + static_cast<size_t>(N) * TC;
    load_tile<TC, NT>(yt, y, g, N, q0);
    load_tile<TC, NT>(kt, k, g, R, q0);
    __syncthreads();
    if (!due)
      return;
    const int nnz = m.nnz_jac;
    const int ne = nnz * TC;
    double* out = jac + g * static_cast<size_t>(nnz) * kGroup + q0 + cc;
    constexpr int B = 2;
    int i = threadIdx.x;
    double v_next = !FROM_ZERO && i < ne ? out[(i / TC) * kGroup] : 0.0;
    for (; i < ne; i += NT)
    {
      const int e = i / TC;
      double v = v_next;
      if (!FROM_ZERO && i + NT < ne)
        v_next = out[((i + NT) / TC) * kGroup];
      const int qe = c.jel_offsets[e + 1];
      for (int q = c.jel_offsets[e]; q < qe; q += B)
      {
        uint64_t pk[B];
        double cf[B], kk[B], y0[B], y1[B];
#pragma unroll
        for (int r = 0; r < B; ++r)
        {
          pk[r] = c.jel_pack[q + r];
          cf[r] = c.jel_coef[q + r];
        }
#pragma unroll
        for (int r = 0; r < B; ++r)
        {
          kk[r] = kt[(pk[r] & 0xffffu) * TC + cc];
          y0[r] = yt[((pk[r] >> 16) & 0xffffu) * TC + cc];
          y1[r] = yt[((pk[r] >> 32) & 0xffffu) * TC + cc];
        }
#pragma unroll
        for (int r = 0; r < B; ++r)
        {
          if (q + r < qe)
          {
            double d;
            if ((pk[r] >> 48) == 0xffffu)
            {
              const int t = c.jel_term[q + r];
              d = kt[m.term_process[t] * TC + cc];
              for (int a = m.term_dep_offsets[t]; a < m.term_dep_offsets[t + 1]; ++a)
                d *= yt[m.term_dep_ids[a] * TC + cc];
            }
            else
            {
              d = packed_product(kk[r], pk[r], y0[r], y1[r], 0.0);
            }
            v += cf[r] * d;
          }
        }
      }
      out[e * kGroup] = v;
    }
  }

  template<int TC, int NT, bool FROM_ZERO = false>
  __global__ void __launch_bounds__(NT) k_subtract_jacobian_tiled(DevMech m, DevCoop c, int64_t n_cells,
                                                                  const double* __restrict__ k,
                                                                  const double* __restrict__ y,
                                                                  double* __restrict__ jac)
  {
    subtract_jacobian_tile<TC, NT, false, FROM_ZERO>(m, c, n_cells, k, y, jac, nullptr);
  }

  template<int TC, int NT>
  __global__ void __launch_bounds__(NT) k_jacobian_due_tiled(DevMech m, DevCoop c, int64_t n_cells,
                                                             const double* __restrict__ k,
                                                             const double* __restrict__ y, double* __restrict__ jac,
                                                             const CellControl* __restrict__ ctl)
  {
    subtract_jacobian_tile<TC, NT, true>(m, c, n_cells, k, y, jac, ctl);
  }

  // dst[map[v] * TC + cc] = src[(g * n_values + v) * 64 + q0 + cc] for map[v] >= 0,
  // eight loads in flight per thread
  template<int TC, int NT>
  __device__ inline void load_tile_scatter(double* dst, const int32_t* map, const double* __restrict__ src, size_t g,
                                           int n_values, int q0)
  {
    const int total = n_values * TC;
    const double* base = src + g * static_cast<size_t>(n_values) * kGroup + q0;
    for (int i0 = threadIdx.x; i0 < total; i0 += 8 * NT)
    {
      double v[8];
      int to[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
      {
        const int i = i0 + j * NT < total ? i0 + j * NT : i0;
        v[j] = base[(i / TC) * kGroup + i % TC];
        to[j] = map[i / TC];
      }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (i0 + j * NT < total && to[j] >= 0)
          dst[to[j] * TC + (i0 + j * NT) % TC] = v[j];
    }
  }

  // Tiled LU and substitution: TC cells of a group per workgroup, their in-place
  // factors in LDS (position p of cell cc at p * TC + cc), the level schedules of the
  // cooperative form run over (task, cell) pairs.  L and U leave in the reference's
  // separate storage.
  template<int TC>
  __global__ void __launch_bounds__(kTileThreads) k_lu_decompose_tiled(DevMech m, DevCoop c, int64_t n_cells,
                                                                       const double* __restrict__ jac,
                                                                       double* __restrict__ lower, double* __restrict__ upper)
  {
    extern __shared__ double alu[];
    size_t g;
    int q0;
    if (!tile_of_block<TC>(n_cells, g, q0))
      return;
    const int tid = threadIdx.x;
    for (int i = tid; i < c.nnz_alu * TC; i += kTileThreads)
      alu[i] = 0.0;
    __syncthreads();
    load_tile_scatter<TC, kTileThreads>(alu, c.jac_to_alu, jac, g, m.nnz_jac, q0);
    __syncthreads();
    const DevLevels& lv = c.lu;
    for (int l = 0; l < lv.n_levels; ++l)
    {
      const int tb = lv.offsets[l], n_items = (lv.offsets[l + 1] - tb) * TC;
      for (int w = tid; w < n_items; w += kTileThreads)
      {
        const int t = tb + w / TC, cc = w % TC;
        const uint2 task = lv.tasks[t];
        const uint32_t qe = lv.tasks[t + 1].y;
        const uint32_t o = task.x & 0xffffu, dg = task.x >> 16;
        double v = task_dot<TC>(alu[o * TC + cc], lv.terms, task.y, qe, alu, alu, cc);
        if (dg < kUnitDivisor)
          v = v / alu[dg * TC + cc];
        alu[o * TC + cc] = v;
      }
      __syncthreads();
    }
    for (int i = tid; i < m.nnz_lower * TC; i += kTileThreads)
    {
      const int p = i / TC, cc = i % TC;
      if (static_cast<int64_t>(g) * kGroup + q0 + cc < n_cells)
      {
        const int a = c.alu_lower[p];
        lower[(g * m.nnz_lower + p) * kGroup + q0 + cc] = a >= 0 ? alu[a * TC + cc] : 1.0;
      }
    }
    for (int i = tid; i < m.nnz_upper * TC; i += kTileThreads)
    {
      const int p = i / TC, cc = i % TC;
      if (static_cast<int64_t>(g) * kGroup + q0 + cc < n_cells)
        upper[(g * m.nnz_upper + p) * kGroup + q0 + cc] = alu[c.alu_upper[p] * TC + cc];
    }
  }

  // x <- U^-1 L^-1 x (linear_solver.inl:235-294), dividing by L[i][i] as the reference does
  template<int TC>
  __global__ void __launch_bounds__(kTileThreads) k_lu_solve_tiled(DevMech m, DevCoop c, int64_t n_cells,
                                                                   const double* __restrict__ lower,
                                                                   const double* __restrict__ upper, double* __restrict__ x,
                                                                   int p_capacity)
  {
    extern __shared__ double lds_solve[];
    const int N = m.n_species;
    double* alu = lds_solve;
    double* xt = alu + static_cast<size_t>(c.nnz_alu) * TC;
    double* ld = xt + static_cast<size_t>(N) * TC;
    size_t g;
    int q0;
    if (!tile_of_block<TC>(n_cells, g, q0))
      return;
    const int tid = threadIdx.x;
    load_tile_scatter<TC, kTileThreads>(alu, c.alu_lower, lower, g, m.nnz_lower, q0);
    load_tile_scatter<TC, kTileThreads>(alu, c.alu_upper, upper, g, m.nnz_upper, q0);
    load_tile<TC, kTileThreads>(xt, x, g, N, q0);
    for (int i = tid; i < N * TC; i += kTileThreads)
    {
      const int r = i / TC, cc = i % TC;
      ld[i] = lower[(g * m.nnz_lower + m.lower_diag[r]) * kGroup + q0 + cc];
    }
    __syncthreads();
    // rows without off-diagonal L entries only divide; they are inputs of every level
    for (int i = tid; i < N * TC; i += kTileThreads)
    {
      const int r = i / TC;
      if (m.fwd_offsets[r] == m.fwd_offsets[r + 1])
        xt[i] = xt[i] / ld[i];
    }
    __syncthreads();
    // one cell per workgroup: the back substitution's long ordered sums run two-phase
    // (products into LDS after ld, then one thread per row), as in the cooperative form
    const int passes = TC == 1 ? 1 : 2;
    for (int pass = 0; pass < passes; ++pass)
    {
      const DevLevels& lv = pass == 0 ? c.fwd : c.bwd;
      for (int l = 0; l < lv.n_levels; ++l)
      {
        const int tb = lv.offsets[l], n_items = (lv.offsets[l + 1] - tb) * TC;
        for (int w = tid; w < n_items; w += kTileThreads)
        {
          const int t = tb + w / TC, cc = w % TC;
          const uint2 task = lv.tasks[t];
          const uint32_t qe = lv.tasks[t + 1].y;
          const uint32_t o = task.x & 0xffffu, dg = task.x >> 16;
          double v = task_dot<TC>(xt[o * TC + cc], lv.terms, task.y, qe, alu, xt, cc);
          if (pass == 0 ? dg == kUnitDivisor : dg < kUnitDivisor)
            v = v / (pass == 0 ? ld[o * TC + cc] : alu[dg * TC + cc]);
          xt[o * TC + cc] = v;
        }
        __syncthreads();
      }
    }
    if constexpr (TC == 1)
      coop_levels<kTileThreads>(c.bwd, c.bwd.tasks, c.bwd.terms, alu, xt, xt, ld + N, p_capacity);
    for (int i = tid; i < N * TC; i += kTileThreads)
    {
      const int r = i / TC, cc = i % TC;
      if (static_cast<int64_t>(g) * kGroup + q0 + cc < n_cells)
        x[(g * N + r) * kGroup + q0 + cc] = xt[i];
    }
  }

  __global__ void __launch_bounds__(kBlock) k_alpha_minus_jacobian(DevMech m, int64_t n_cells, double* __restrict__ jac,
                                                                   double alpha)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    Lane J = lane_view(jac, c / kGroup, m.nnz_jac, c % kGroup);
    for (int i = 0; i < m.n_species; ++i)
      J[m.jac_diag[i]] += alpha * 1.0;
  }

  __global__ void __launch_bounds__(kBlock) k_lu_decompose(DevMech m, int64_t n_cells, const double* __restrict__ jac,
                                                           double* __restrict__ lower, double* __restrict__ upper)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    const size_t g = c / kGroup;
    const int l = c % kGroup;
    lu_decompose(m, lane_view(const_cast<double*>(jac), g, m.nnz_jac, l), lane_view(lower, g, m.nnz_lower, l),
                 lane_view(upper, g, m.nnz_upper, l));
  }

  __global__ void __launch_bounds__(kBlock) k_lu_solve(DevMech m, int64_t n_cells, const double* __restrict__ lower,
                                                       const double* __restrict__ upper, double* __restrict__ x)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    const size_t g = c / kGroup;
    const int l = c % kGroup;
    lu_solve<false>(m, lane_view(const_cast<double*>(lower), g, m.nnz_lower, l),
                    lane_view(const_cast<double*>(upper), g, m.nnz_upper, l), lane_view(x, g, m.n_species, l));
  }

  __global__ void __launch_bounds__(kBlock) k_normalized_error(int n_species, int64_t n_cells, const double* __restrict__ y,
                                                               const double* __restrict__ ynew,
                                                               const double* __restrict__ yerr,
                                                               const double* __restrict__ atol, double rtol,
                                                               double* __restrict__ err)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    const size_t g = c / kGroup;
    const int l = c % kGroup;
    const double s = error_sum(n_species, lane_view(const_cast<double*>(y), g, n_species, l),
                               lane_view(const_cast<double*>(ynew), g, n_species, l),
                               lane_view(const_cast<double*>(yerr), g, n_species, l), atol, rtol);
    err[c] = smax(sqrt(s / static_cast<double>(n_species > 0 ? n_species : 1)), 1.0e-10);
  }

  // ------------------------------------------------------------------ solver state

  struct Work
  {
    double* y;       // N
    double* k;       // R
    double* jac;     // nnzJ
    double* lower;   // nnzL
    double* upper;   // nnzU
    double* f0;      // N
    double* ynew;    // N
    double* yerr;    // N
    double* K;       // stage vectors, N each, stage j at K + j * k_stride
    size_t k_stride;
  };

  // The stage vectors of one lane: stage j starts k_stride elements after stage j-1.
  struct Stages
  {
    double* p;
    size_t stride;
    __device__ __forceinline__ Lane operator[](int j) const
    {
      return Lane{ p + static_cast<size_t>(j) * stride };
    }
  };

  __device__ __forceinline__ Stages stage_view(const Work& w, size_t group, int n, int lane)
  {
    return Stages{ lane_view(w.K, group, n, lane).p, w.k_stride };
  }

  __device__ __forceinline__ void load_control(const CellControl* ctl, int64_t c, CellControl& out)
  {
    out = ctl[c];
  }

  // Stage s of one attempt for one lane up to its linear solve: K_s holds the
  // right-hand side (rosenbrock.inl:140-191).  PART 1 stops before the forcing of a
  // new function evaluation (ynew formed, K_s zeroed), PART 2 is the rest after it,
  // PART 3 both with the forcing in between.
  template<int PART = 3>
  __device__ inline void stage_rhs(const DevMech& m, const DevParams& p, int s, const Lane& y, const Lane& k,
                                   const Lane& f0, const Lane& ynew, const Stages& K, CellControl& c)
  {
    const int n = m.n_species;
    const int sc = s * (s - 1) / 2;
    const Lane Ks = K[s];
    if (s == 0)
    {
      if constexpr ((PART & 2) != 0)
        for (int i = 0; i < n; ++i)
          Ks[i] = f0[i];
    }
    else if (p.new_f[s])
    {
      if constexpr ((PART & 1) != 0)
      {
        for (int i = 0; i < n; ++i)
        {
          double v = y[i];
          for (int j = 0; j < s; ++j)
            v += p.a[sc + j] * K[j][i];
          ynew[i] = v;
        }
        for (int i = 0; i < n; ++i)
          Ks[i] = 0.0;
        c.stats[MICM_HIP_STAT_FUNCTION_CALLS] += 1;
      }
      if constexpr (PART == 3)
        add_forcing(m, k, ynew, Ks);
    }
    if constexpr ((PART & 2) == 0)
      return;
    if (s + 1 < p.stages && !p.new_f[s + 1])
    {
      const Lane Kn = K[s + 1];
      for (int i = 0; i < n; ++i)
        Kn[i] = Ks[i];
    }
    if (s > 0)
    {
      double coef[6];
      for (int j = 0; j < s; ++j)
        coef[j] = p.c[sc + j] / c.H;
      for (int i = 0; i < n; ++i)
      {
        double v = Ks[i];
        for (int j = 0; j < s; ++j)
          v += coef[j] * K[j][i];
        Ks[i] = v;
      }
    }
    c.stats[MICM_HIP_STAT_SOLVES] += 1;
  }

  // Stage s of one attempt for one lane (rosenbrock.inl:140-195).
  __device__ inline void run_stage(const DevMech& m, const DevParams& p, int s, const Lane& y, const Lane& k,
                                   const Lane& f0, const Lane& ynew, const Stages& K, const Lane& L, const Lane& U,
                                   CellControl& c)
  {
    stage_rhs(m, p, s, y, k, f0, ynew, K, c);
    lu_solve<true>(m, L, U, K[s]);
  }

  // Ynew = Y + sum m_s K_s, Yerr = sum e_s K_s; returns the error sum over species.
  __device__ inline double combine_and_error(const DevMech& m, const DevParams& p, const Lane& y, const Lane& ynew,
                                             const Lane& yerr, const Stages& K, const double* atol)
  {
    const int n = m.n_species;
    for (int i = 0; i < n; ++i)
    {
      double v = y[i], e = 0.0;
      for (int s = 0; s < p.stages; ++s)
        v += p.m[s] * K[s][i];
      for (int s = 0; s < p.stages; ++s)
        e += p.e[s] * K[s][i];
      ynew[i] = v;
      yerr[i] = e;
    }
    return error_sum(n, y, ynew, yerr, atol, p.rtol);
  }

  // Start of an attempt: a new step (forcing + Jacobian) if due, then alpha and LU.
  // Returns false if the cell finished at the loop-top checks.
  // FORCING false: f(y) is left to the caller (f0 zeroed here).
  template<bool FORCING = true>
  __device__ inline bool prepare_attempt(const DevMech& m, const DevParams& p, const Lane& y, const Lane& k,
                                         const Lane& f0, const Lane& J, CellControl& c)
  {
    if (c.flags & kNewStep)
    {
      if (!begin_step(p, c))
        return false;
      for (int i = 0; i < m.n_species; ++i)
        f0[i] = 0.0;
      if constexpr (FORCING)
        add_forcing(m, k, y, f0);
      c.stats[MICM_HIP_STAT_FUNCTION_CALLS] += 1;
      for (int i = 0; i < m.nnz_jac; ++i)
        J[i] = 0.0;
      subtract_jacobian(m, k, y, J);
      c.stats[MICM_HIP_STAT_JACOBIAN_UPDATES] += 1;
      c.last_alpha = 0.0;
      c.flags &= ~kNewStep;
    }
    const double alpha = next_alpha(p, c);
    for (int i = 0; i < m.n_species; ++i)
      J[m.jac_diag[i]] += alpha * 1.0;
    c.stats[MICM_HIP_STAT_DECOMPOSITIONS] += 1;
    return true;
  }

  __device__ inline bool factor_attempt(const DevMech& m, const DevParams& p, const Lane& y, const Lane& k,
                                        const Lane& f0, const Lane& J, const Lane& L, const Lane& U, CellControl& c)
  {
    if (!prepare_attempt(m, p, y, k, f0, J, c))
      return false;
    lu_decompose(m, J, L, U);
    return true;
  }

  // ---- plain form: one kernel per phase over all cells

  __global__ void __launch_bounds__(kBlock) k_plain_init(DevParams p, int64_t n_cells, CellControl* ctl)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    CellControl cc;
    control_init(p, cc);
    ctl[c] = cc;
  }

  __global__ void __launch_bounds__(kBlock) k_plain_factor(DevMech m, DevParams p, int64_t n_cells, Work w, CellControl* ctl)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    CellControl cc = ctl[c];
    if (cc.status != MICM_HIP_STATE_RUNNING)
      return;
    const size_t g = c / kGroup;
    const int l = c % kGroup;
    factor_attempt(m, p, lane_view(w.y, g, m.n_species, l), lane_view(w.k, g, m.n_reactions, l),
                   lane_view(w.f0, g, m.n_species, l), lane_view(w.jac, g, m.nnz_jac, l),
                   lane_view(w.lower, g, m.nnz_lower, l), lane_view(w.upper, g, m.nnz_upper, l), cc);
    ctl[c] = cc;
  }

  // With the level schedules available the forcing, the factorisation and the
  // substitutions run as the tiled kernels over every cell instead (finished cells'
  // values are not used); k_plain_prepare (or k_plain_begin, the masked tiled
  // Jacobian and k_plain_alpha) and k_stage_rhs_elem do the per-cell rest of
  // k_plain_factor and k_plain_stage.  f(y) is evaluated for every running cell at
  // every attempt: y is unchanged since the cell's step began, so it is the same value
  // the step started with.
  __global__ void __launch_bounds__(kBlock) k_plain_prepare(DevMech m, DevParams p, int64_t n_cells, Work w,
                                                            CellControl* ctl)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    CellControl cc = ctl[c];
    if (cc.status != MICM_HIP_STATE_RUNNING)
      return;
    const size_t g = c / kGroup;
    const int l = c % kGroup;
    const Lane f0 = lane_view(w.f0, g, m.n_species, l);
    for (int i = 0; i < m.n_species; ++i)
      f0[i] = 0.0;
    prepare_attempt<false>(m, p, lane_view(w.y, g, m.n_species, l), lane_view(w.k, g, m.n_reactions, l), f0,
                           lane_view(w.jac, g, m.nnz_jac, l), cc);
    ctl[c] = cc;
  }

  // The same start of an attempt with the Jacobian formed by k_jacobian_due_tiled
  // between the two halves: k_plain_begin opens a new step where one is due (and
  // marks the cell kJacobianDue), k_plain_alpha adds alpha to the diagonal.
  __global__ void __launch_bounds__(kBlock) k_plain_begin(DevMech m, DevParams p, int64_t n_cells, Work w,
                                                          CellControl* ctl)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    CellControl cc = ctl[c];
    if (cc.status != MICM_HIP_STATE_RUNNING)
      return;
    if (cc.flags & kNewStep)
    {
      if (begin_step(p, cc))
      {
        cc.stats[MICM_HIP_STAT_FUNCTION_CALLS] += 1;
        cc.stats[MICM_HIP_STAT_JACOBIAN_UPDATES] += 1;
        cc.last_alpha = 0.0;
        cc.flags = (cc.flags & ~kNewStep) | kJacobianDue;
      }
    }
    ctl[c] = cc;
  }

  __global__ void __launch_bounds__(kBlock) k_plain_alpha(DevMech m, DevParams p, int64_t n_cells, Work w,
                                                          CellControl* ctl)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    CellControl cc = ctl[c];
    if (cc.status != MICM_HIP_STATE_RUNNING)
      return;
    const Lane J = lane_view(w.jac, c / kGroup, m.nnz_jac, c % kGroup);
    const double alpha = next_alpha(p, cc);
    for (int i = 0; i < m.n_species; ++i)
      J[m.jac_diag[i]] += alpha * 1.0;
    cc.stats[MICM_HIP_STAT_DECOMPOSITIONS] += 1;
    cc.flags &= ~kJacobianDue;
    ctl[c] = cc;
  }

  // stage_rhs<PART> with one thread per (cell, species) of the running cells, in the
  // grouped layout's order: the same expressions element by element; the thread of
  // species 0 counts the cell's function call or solve.
  template<int PART>
  __global__ void __launch_bounds__(kBlock) k_stage_rhs_elem(DevMech m, DevParams p, int64_t n_cells, Work w,
                                                             CellControl* ctl, int s)
  {
    const int n = m.n_species;
    const int64_t per_group = static_cast<int64_t>(n) * kGroup;
    const int64_t f = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (f >= n_groups_of(n_cells) * per_group)
      return;
    const int64_t rem = f % per_group;
    const int64_t c = (f / per_group) * kGroup + rem % kGroup;
    if (c >= n_cells || ctl[c].status != MICM_HIP_STATE_RUNNING)
      return;
    const bool first = rem < kGroup;  // species 0 of the cell
    const int sc = s * (s - 1) / 2;
    double* K = w.K + f;
    double* Ks = K + static_cast<size_t>(s) * w.k_stride;
    if (s == 0)
    {
      if constexpr ((PART & 2) != 0)
        *Ks = w.f0[f];
    }
    else if (p.new_f[s])
    {
      if constexpr ((PART & 1) != 0)
      {
        double v = w.y[f];
        for (int j = 0; j < s; ++j)
          v += p.a[sc + j] * K[static_cast<size_t>(j) * w.k_stride];
        w.ynew[f] = v;
        if (first)
          ctl[c].stats[MICM_HIP_STAT_FUNCTION_CALLS] += 1;
      }
    }
    if constexpr ((PART & 2) == 0)
      return;
    if (s + 1 < p.stages && !p.new_f[s + 1])
      Ks[w.k_stride] = *Ks;
    if (s > 0)
    {
      const double H = ctl[c].H;
      double v = *Ks;
      for (int j = 0; j < s; ++j)
        v += (p.c[sc + j] / H) * K[static_cast<size_t>(j) * w.k_stride];
      *Ks = v;
    }
    if (first)
      ctl[c].stats[MICM_HIP_STAT_SOLVES] += 1;
  }

  __global__ void __launch_bounds__(kBlock) k_plain_stage(DevMech m, DevParams p, int64_t n_cells, Work w, CellControl* ctl,
                                                          int s)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    CellControl cc = ctl[c];
    if (cc.status != MICM_HIP_STATE_RUNNING)
      return;
    const size_t g = c / kGroup;
    const int l = c % kGroup;
    const int n = m.n_species;
    const Stages K = stage_view(w, g, n, l);
    run_stage(m, p, s, lane_view(w.y, g, n, l), lane_view(w.k, g, m.n_reactions, l), lane_view(w.f0, g, n, l),
              lane_view(w.ynew, g, n, l), K, lane_view(w.lower, g, m.nnz_lower, l), lane_view(w.upper, g, m.nnz_upper, l),
              cc);
    ctl[c].stats[MICM_HIP_STAT_FUNCTION_CALLS] = cc.stats[MICM_HIP_STAT_FUNCTION_CALLS];
    ctl[c].stats[MICM_HIP_STAT_SOLVES] = cc.stats[MICM_HIP_STAT_SOLVES];
  }

  // Error and accept/reject of the per-cell step control, after k_combine_elem.
  __global__ void __launch_bounds__(kBlock) k_plain_finish(DevMech m, DevParams p, int64_t n_cells, Work w, CellControl* ctl,
                                                           const double* __restrict__ atol, unsigned int* active)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    bool running = false;
    if (c < n_cells)
    {
      CellControl cc = ctl[c];
      if (cc.status == MICM_HIP_STATE_RUNNING)
      {
        const size_t g = c / kGroup;
        const int l = c % kGroup;
        const int n = m.n_species;
        Lane y = lane_view(w.y, g, n, l), ynew = lane_view(w.ynew, g, n, l);
        const double s = error_sum(n, y, ynew, lane_view(w.yerr, g, n, l), atol, p.rtol);
        const double error = smax(sqrt(s / static_cast<double>(n > 0 ? n : 1)), 1.0e-10);
        if (control_update(p, cc, error))
          for (int i = 0; i < n; ++i)
            y[i] = ynew[i];
        ctl[c] = cc;
        running = cc.status == MICM_HIP_STATE_RUNNING;
      }
    }
    const unsigned long long mask = __ballot(running);
    if ((threadIdx.x % kGroup) == 0 && mask)
      atomicAdd(active, static_cast<unsigned int>(__popcll(mask)));
  }

  // Ynew = Y + sum m_s K_s and Yerr = sum e_s K_s (combine_and_error's expressions)
  // with one thread per (cell, species) of the running cells, in the grouped layout's
  // order; k_plain_finish or (lockstep) k_lockstep_error and k_lockstep_control
  // finish the attempt.
  __global__ void __launch_bounds__(kBlock) k_combine_elem(DevMech m, DevParams p, int64_t n_cells, Work w,
                                                               const CellControl* __restrict__ ctl)
  {
    const int64_t per_group = static_cast<int64_t>(m.n_species) * kGroup;
    const int64_t f = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (f >= n_groups_of(n_cells) * per_group)
      return;
    const int64_t c = (f / per_group) * kGroup + f % kGroup;
    if (c >= n_cells || ctl[c].status != MICM_HIP_STATE_RUNNING)
      return;
    double v = w.y[f], e = 0.0;
    for (int s = 0; s < p.stages; ++s)
      v += p.m[s] * w.K[static_cast<size_t>(s) * w.k_stride + f];
    for (int s = 0; s < p.stages; ++s)
      e += p.e[s] * w.K[static_cast<size_t>(s) * w.k_stride + f];
    w.ynew[f] = v;
    w.yerr[f] = e;
  }

  constexpr int kLockstepErrorThreads = 1024;

  // The reference's NormalizedError for one State (rosenbrock.inl:353-402): one running
  // sum of the squared scaled errors in the order the State's dense matrix policy visits
  // them, then max(sqrt(sum / (cells * species)), 1e-10).  `group` is that policy's
  // vector length: 1 for Matrix<Real> (cell by cell, species inner), L for
  // VectorMatrix<L> (each group of L cells species by species, the group's cells inner,
  // a partial last group over its own cells; vector_matrix.hpp:860-867).  The terms are
  // formed by the whole workgroup a block at a time and added one after another by
  // thread 0.  Every thread of the workgroup calls it; the result is thread 0's.
  __device__ inline double ordered_error(int n, int64_t n_cells, int64_t group, const double* __restrict__ y,
                                         const double* __restrict__ ynew, const double* __restrict__ yerr,
                                         const double* __restrict__ atol, double rtol, double* terms)
  {
    const int64_t n_terms = n_cells * n;
    const int64_t per_group = group * n;
    double acc = 0.0;
    for (int64_t base = 0; base < n_terms; base += kLockstepErrorThreads)
    {
      const int64_t f = base + threadIdx.x;
      double q = 0.0;
      if (f < n_terms)
      {
        const int64_t gi = f / per_group;
        const int64_t off = f - gi * per_group;
        const int64_t rows = n_cells - gi * group < group ? n_cells - gi * group : group;
        const int i = static_cast<int>(off / rows);
        const int64_t c = gi * group + off % rows;
        const size_t g = c / kGroup;
        const int cl = c % kGroup;
        const double a = fabs(lane_view(const_cast<double*>(y), g, n, cl)[i]);
        const double b = fabs(lane_view(const_cast<double*>(ynew), g, n, cl)[i]);
        const double ymax = a > b ? a : b;
        const double v = lane_view(const_cast<double*>(yerr), g, n, cl)[i] / (atol[i] + rtol * ymax);
        q = v * v;
      }
      terms[threadIdx.x] = q;
      __syncthreads();
      if (threadIdx.x == 0)
      {
        const int cnt = static_cast<int>(n_terms - base < kLockstepErrorThreads ? n_terms - base
                                                                                : kLockstepErrorThreads);
        // sixteen reads in flight, then their additions in order
        int j = 0;
        for (; j + 16 <= cnt; j += 16)
        {
          double t[16];
#pragma unroll
          for (int u = 0; u < 16; ++u)
            t[u] = terms[j + u];
#pragma unroll
          for (int u = 0; u < 16; ++u)
            acc += t[u];
        }
        for (; j < cnt; ++j)
          acc += terms[j];
      }
      __syncthreads();
    }
    const double N = static_cast<double>(n_terms > 0 ? n_terms : 1);
    return smax(sqrt(acc / N), 1.0e-10);
  }

  // Lockstep error of the state's attempt, in the order of the state's dense matrix policy.
  __device__ inline double lockstep_error(const DevMech& m, const DevParams& p, int64_t n_cells, const Work& w,
                                          const double* __restrict__ atol, double* terms)
  {
    return ordered_error(m.n_species, n_cells, p.error_group, w.y, w.ynew, w.yerr, atol, p.rtol, terms);
  }

  __global__ void __launch_bounds__(kLockstepErrorThreads) k_state_error(int n, int64_t n_cells, int64_t group,
                                                                         const double* y, const double* ynew,
                                                                         const double* yerr, const double* atol,
                                                                         double rtol, double* error_out)
  {
    __shared__ double terms[kLockstepErrorThreads];
    const double e = ordered_error(n, n_cells, group, y, ynew, yerr, atol, rtol, terms);
    if (threadIdx.x == 0)
      *error_out = e;
  }

  __global__ void __launch_bounds__(kLockstepErrorThreads) k_lockstep_error(DevMech m, DevParams p, int64_t n_cells,
                                                                            Work w, const CellControl* ctl,
                                                                            const double* __restrict__ atol,
                                                                            double* error_out)
  {
    __shared__ double terms[kLockstepErrorThreads];
    // the cells share one control state: all stopped at the loop-top checks or none
    if (ctl[0].status != MICM_HIP_STATE_RUNNING)
      return;
    const double e = lockstep_error(m, p, n_cells, w, atol, terms);
    if (threadIdx.x == 0)
      *error_out = e;
  }

  __global__ void __launch_bounds__(kBlock) k_lockstep_control(DevMech m, DevParams p, int64_t n_cells, Work w,
                                                               CellControl* ctl, const double* error, unsigned int* active)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    bool running = false;
    if (c < n_cells)
    {
      CellControl cc = ctl[c];
      if (cc.status == MICM_HIP_STATE_RUNNING)
      {
        const size_t g = c / kGroup;
        const int l = c % kGroup;
        Lane y = lane_view(w.y, g, m.n_species, l), ynew = lane_view(w.ynew, g, m.n_species, l);
        if (control_update(p, cc, *error))
          for (int i = 0; i < m.n_species; ++i)
            y[i] = ynew[i];
        ctl[c] = cc;
        running = cc.status == MICM_HIP_STATE_RUNNING;
      }
    }
    const unsigned long long mask = __ballot(running);
    if ((threadIdx.x % kGroup) == 0 && mask)
      atomicAdd(active, static_cast<unsigned int>(__popcll(mask)));
  }

  // Lockstep loop-top: all cells share one control state, so a cell that stops at the
  // loop-top checks stops them all; the factor kernel handles it per cell identically.

  // Lockstep form of a small State in one workgroup: per attempt, every cell's
  // factorisation, stages and combination (k_plain_factor, k_plain_stage,
  // k_combine_elem), the State's error sum in the reference's order
  // (k_lockstep_error) and the shared step control (k_lockstep_control),
  // separated by workgroup barriers instead of kernel launches and a host check per
  // attempt.  Same operations in the same order, so the results are those of the
  // launch-per-phase loop.  Every exit condition and every barrier is uniform over
  // the workgroup.
  constexpr int kFusedLockstepThreads = kLockstepErrorThreads;
  constexpr int64_t kFusedLockstepMaxCells = kFusedLockstepThreads;
  constexpr int64_t kFusedLockstepMaxFactor = 256;  // nnz of the in-place factor

  __global__ void __launch_bounds__(kFusedLockstepThreads) k_lockstep_fused(DevMech m, DevParams p, int64_t n_cells,
                                                                            Work w, CellControl* ctl,
                                                                            const double* __restrict__ atol,
                                                                            int64_t max_iterations)
  {
    __shared__ double s_error;
    __shared__ double s_terms[kFusedLockstepThreads];
    const int n = m.n_species;
    for (int64_t it = 0; it < max_iterations; ++it)
    {
      for (int64_t c = threadIdx.x; c < n_cells; c += kFusedLockstepThreads)
      {
        CellControl cc = ctl[c];
        if (cc.status != MICM_HIP_STATE_RUNNING)
          continue;
        const size_t g = c / kGroup;
        const int l = c % kGroup;
        const Lane y = lane_view(w.y, g, n, l), k = lane_view(w.k, g, m.n_reactions, l);
        const Lane f0 = lane_view(w.f0, g, n, l), ynew = lane_view(w.ynew, g, n, l);
        const Lane L = lane_view(w.lower, g, m.nnz_lower, l), U = lane_view(w.upper, g, m.nnz_upper, l);
        const Stages K = stage_view(w, g, n, l);
        if (factor_attempt(m, p, y, k, f0, lane_view(w.jac, g, m.nnz_jac, l), L, U, cc))
        {
          for (int s = 0; s < p.stages; ++s)
            run_stage(m, p, s, y, k, f0, ynew, K, L, U, cc);
          combine_and_error(m, p, y, ynew, lane_view(w.yerr, g, n, l), K, atol);
        }
        ctl[c] = cc;
      }
      __syncthreads();
      // the cells share one control state: all stopped at the loop-top checks or none
      if (ctl[0].status != MICM_HIP_STATE_RUNNING)
        break;
      {
        const double e = lockstep_error(m, p, n_cells, w, atol, s_terms);
        if (threadIdx.x == 0)
          s_error = e;
      }
      __syncthreads();
      const double error = s_error;
      int running = 0;
      for (int64_t c = threadIdx.x; c < n_cells; c += kFusedLockstepThreads)
      {
        CellControl cc = ctl[c];
        if (cc.status != MICM_HIP_STATE_RUNNING)
          continue;
        const size_t g = c / kGroup;
        const int l = c % kGroup;
        Lane y = lane_view(w.y, g, n, l), ynew = lane_view(w.ynew, g, n, l);
        if (control_update(p, cc, error))
          for (int i = 0; i < n; ++i)
            y[i] = ynew[i];
        ctl[c] = cc;
        running |= cc.status == MICM_HIP_STATE_RUNNING ? 1 : 0;
      }
      if (!__syncthreads_or(running))
        break;
    }
  }

  // Post-solve clamp (solver.hpp:275-297) and per-cell results.
  __global__ void __launch_bounds__(kBlock) k_finalize(int n_species, int64_t n_cells, double* y, CellControl* ctl,
                                                       int32_t* status, int64_t* stats, double* final_time)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_cells)
      return;
    CellControl cc = ctl[c];
    if (cc.status == MICM_HIP_STATE_RUNNING)
      cc.status = MICM_HIP_STATE_CONVERGED;
    status[c] = cc.status;
    for (int i = 0; i < 7; ++i)
      stats[c * 7 + i] = cc.stats[i];
    final_time[c] = cc.t;
    Lane yy = lane_view(y, c / kGroup, n_species, c % kGroup);
    for (int i = 0; i < n_species; ++i)
      yy[i] = smax(yy[i], 0.0);
  }

  __global__ void k_zero_u32(unsigned int* p)
  {
    *p = 0u;
  }

  __global__ void k_zero_i32(int* p)
  {
    *p = 0;
  }

  // ---- processing order of the per-cell forms: the cells of the previous Solve,
  // most step attempts first, so that the longest cells do not start last.  A
  // counting sort over 64 bins of the attempt count (four bins per power of two);
  // any permutation gives the same results, so ties may land in any order.
  constexpr int kOrderBins = 64;

  __device__ inline int order_bin(int64_t attempts)
  {
    const uint64_t a = attempts > 1 ? static_cast<uint64_t>(attempts) : 1u;
    const int msb = 63 - __clzll(a);
    const int frac = msb >= 2 ? static_cast<int>((a >> (msb - 2)) & 3u) : 0;
    const int b = msb * 4 + frac;
    return kOrderBins - 1 - (b < kOrderBins - 1 ? b : kOrderBins - 1);  // heaviest first
  }

  __global__ void __launch_bounds__(kBlock) k_order_clear(int* bins)
  {
    if (threadIdx.x <= kOrderBins)
      bins[threadIdx.x] = 0;
  }

  __global__ void __launch_bounds__(kBlock) k_order_count(const int64_t* __restrict__ stats, int64_t n_cells, int* bins)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c < n_cells)
      atomicAdd(&bins[order_bin(stats[c * 7 + MICM_HIP_STAT_STEPS])], 1);
  }

  // exclusive prefix sum of the bins; bins[kOrderBins] = number of cells
  __global__ void k_order_offsets(int* bins)
  {
    if (threadIdx.x == 0)
    {
      int acc = 0;
      for (int b = 0; b < kOrderBins; ++b)
      {
        const int v = bins[b];
        bins[b] = acc;
        acc += v;
      }
      bins[kOrderBins] = acc;
    }
  }

  __global__ void __launch_bounds__(kBlock) k_order_place(const int64_t* __restrict__ stats, int64_t n_cells, int* bins,
                                                          int32_t* order)
  {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c < n_cells)
      order[atomicAdd(&bins[order_bin(stats[c * 7 + MICM_HIP_STAT_STEPS])], 1)] = static_cast<int32_t>(c);
  }

  // ---- persistent form

  struct Queue
  {
    int* next;  // next unclaimed position in `order`
    const int32_t* order;  // cell processing order (may be null: identity)
  };

  // Each wavefront owns one scratch slot (grouped layout, group = slot).  Lanes take
  // cells from the queue, copy the cell's rate constants and concentrations into the
  // slot, and run attempts until the cell finishes; the result is clamped and written
  // back, and the lane takes the next cell.  The loop ends when the queue is drained
  // and every lane is idle, so every wavefront exits.
  __global__ void __launch_bounds__(kBlock) k_persistent(DevMech m, DevParams p, int64_t n_cells, Queue q, Work slot,
                                                         double* __restrict__ y_state, double* __restrict__ k_state,
                                                         const double* __restrict__ atol, int32_t* status_out,
                                                         int64_t* stats_out, double* final_time_out,
                                                         CellControl* handoff_control, int32_t* handoff_list,
                                                         int* handoff_count, const double* __restrict__ cond_state,
                                                         const double* __restrict__ custom_state)
  {
    const int l = threadIdx.x % kGroup;
    const size_t s = static_cast<size_t>(blockIdx.x) * (kBlock / kGroup) + threadIdx.x / kGroup;
    const int n = m.n_species;
    Lane y = lane_view(slot.y, s, n, l), k = lane_view(slot.k, s, m.n_reactions, l);
    Lane f0 = lane_view(slot.f0, s, n, l), ynew = lane_view(slot.ynew, s, n, l), yerr = lane_view(slot.yerr, s, n, l);
    Lane J = lane_view(slot.jac, s, m.nnz_jac, l), L = lane_view(slot.lower, s, m.nnz_lower, l),
         U = lane_view(slot.upper, s, m.nnz_upper, l);
    const Stages K = stage_view(slot, s, n, l);

    int64_t cell = -1;
    bool drained = false;
    CellControl cc;
    while (true)
    {
      // refill idle lanes with one atomic per wavefront
      const bool want = cell < 0 && !drained;
      const unsigned long long mask = __ballot(want);
      if (mask)
      {
        const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        int base = 0;
        if (l == leader)
          base = atomicAdd(q.next, __popcll(mask));
        base = __shfl(base, leader);
        if (want)
        {
          const int pos = base + __popcll(mask & ((1ull << l) - 1ull));
          if (pos < n_cells)
          {
            cell = q.order ? q.order[pos] : pos;
            Lane ys = lane_view(y_state, cell / kGroup, n, cell % kGroup);
            Lane ks = lane_view(k_state, cell / kGroup, m.n_reactions, cell % kGroup);
            for (int i = 0; i < n; ++i)
              y[i] = ys[i];
            if (cond_state)
            {
              // UpdateStateParameters fused into the launch
              Lane cd = lane_view(const_cast<double*>(cond_state), cell / kGroup, 3, cell % kGroup);
              Lane cu = lane_view(const_cast<double*>(custom_state), cell / kGroup, m.n_custom, cell % kGroup);
              const double T = cd[0], P = cd[1], rho = cd[2];
              for (int r = 0; r < m.n_reactions; ++r)
              {
                const double kr = rate_constant(m, r, T, P, rho, cu);
                k[r] = kr;
                ks[r] = kr;
              }
            }
            else
            {
              for (int r = 0; r < m.n_reactions; ++r)
                k[r] = ks[r];
            }
            control_init(p, cc);
          }
          else
            drained = true;
        }
      }
      if (!__any(cell >= 0))
        break;
      if (cell >= 0 && p.handoff_steps > 0 && (cc.flags & kNewStep) && cc.stats[MICM_HIP_STAT_STEPS] >= p.handoff_steps)
      {
        // a slow cell: park it at this step boundary for the cooperative form
        Lane ys = lane_view(y_state, cell / kGroup, n, cell % kGroup);
        for (int i = 0; i < n; ++i)
          ys[i] = y[i];
        handoff_control[cell] = cc;
        handoff_list[atomicAdd(handoff_count, 1)] = static_cast<int32_t>(cell);
        cell = -1;
      }
      if (cell >= 0)
      {
        bool running = factor_attempt(m, p, y, k, f0, J, L, U, cc);
        if (running)
        {
          for (int st = 0; st < p.stages; ++st)
            run_stage(m, p, st, y, k, f0, ynew, K, L, U, cc);
          const double e = combine_and_error(m, p, y, ynew, yerr, K, atol);
          const double error = smax(sqrt(e / static_cast<double>(n > 0 ? n : 1)), 1.0e-10);
          if (control_update(p, cc, error))
            for (int i = 0; i < n; ++i)
              y[i] = ynew[i];
          running = cc.status == MICM_HIP_STATE_RUNNING;
        }
        if (!running)
        {
          Lane ys = lane_view(y_state, cell / kGroup, n, cell % kGroup);
          for (int i = 0; i < n; ++i)
            ys[i] = smax(y[i], 0.0);
          status_out[cell] = cc.status;
          for (int i = 0; i < 7; ++i)
            stats_out[cell * 7 + i] = cc.stats[i];
          final_time_out[cell] = cc.t;
          cell = -1;
        }
      }
    }
  }

  // ---- cooperative form: one workgroup per cell, working set in LDS

  // STAGE: bit 0 the substitution tables, bit 1 the LU tables, copied into LDS once per
  // workgroup after the per-cell data.  A compile-time choice, so that every table
  // access is a plain LDS or global access.  COMPACT (with STAGE 0 and at least one
  // thread per species): the compact layout of coop_layout_compact, with k_wg holding
  // gridDim.x rate-constant rows of n_reactions doubles.
  //
  // With 256 threads per cell the kernel is compiled for three waves per SIMD (at most
  // 168 VGPRs, a few of them spilled), which lets a third cell onto a CU where LDS
  // allows (TS1: 1.26x); the compact layout also has builds for two waves per SIMD (256
  // registers) and without a bound (MINW 2, 1): where LDS allows two cells per CU anyway
  // (210/560) the one that keeps two cells per CU with the fewest spills is taken.
  template<int NT>
  constexpr int coop_min_waves_per_eu()
  {
    return NT == 256 ? 3 : 1;
  }

  // RECORD: a recorded Solve, appending each cell's checkpoints to `tape` (a separate
  // build, so that the ordinary Solve's register allocation does not carry it)
  template<int NT, int STAGE, bool COMPACT = false, int MINW = coop_min_waves_per_eu<NT>(), bool RECORD = false>
  __global__ void __launch_bounds__(NT, MINW) k_coop(DevMech m, DevCoop c, DevParams p, int p_capacity, int64_t n_fixed,
                                               const int32_t* __restrict__ list, const int* list_count, int* next,
                                               const CellControl* resume, double* __restrict__ y_state,
                                               double* __restrict__ k_state, const double* __restrict__ atol,
                                               int32_t* status_out, int64_t* stats_out, double* final_time_out,
                                               const double* __restrict__ cond_state,
                                               const double* __restrict__ custom_state, double* __restrict__ k_wg,
                                               Tape tape)
  {
    static_assert(!COMPACT || STAGE == 0, "the compact layout reads the tables from global memory");
    extern __shared__ double lds[];
    __shared__ int s_idx;
    __shared__ double s_error;
    const int lane = threadIdx.x;
    const int n = m.n_species;
    const CoopLds L = COMPACT ? coop_layout_compact(lds, m, c, p.stages, p_capacity,
                                                    k_wg + static_cast<size_t>(blockIdx.x) * m.n_reactions)
                              : coop_layout(lds, m, c, p.stages, p_capacity);
    // compact layout: this thread's species' f(y), accumulated Jacobian diagonal and
    // new value (threadIdx.x < n)
    double f0_r = 0.0, jdiag_r = 0.0, ynew_r = 0.0;
    const int64_t count = list ? static_cast<int64_t>(*list_count) : n_fixed;
    CoopTables T{ c.lu.tasks, c.fwd.tasks, c.bwd.tasks, c.lu.terms, c.fwd.terms, c.bwd.terms };
    {
      uint32_t* w = reinterpret_cast<uint32_t*>(lds + coop_lds_doubles(c.nnz_alu, m.nnz_jac, n, p.stages,
                                                                         m.n_reactions, c.n_terms, p_capacity));
      auto stage_copy = [&](const uint32_t* src, size_t words) {
        for (size_t q = lane; q < words; q += NT)
          w[q] = src[q];
        const uint32_t* dst = w;
        w += (words + 1) & ~size_t(1);  // keep 8-byte alignment for the task pairs
        return dst;
      };
      if constexpr ((STAGE & 1) != 0)
      {
        T.fwd_tasks = reinterpret_cast<const uint2*>(
            stage_copy(reinterpret_cast<const uint32_t*>(c.fwd.tasks), 2 * (static_cast<size_t>(c.fwd.n_tasks) + 1)));
        T.fwd_terms = stage_copy(c.fwd.terms, c.fwd.n_terms);
        T.bwd_tasks = reinterpret_cast<const uint2*>(
            stage_copy(reinterpret_cast<const uint32_t*>(c.bwd.tasks), 2 * (static_cast<size_t>(c.bwd.n_tasks) + 1)));
        T.bwd_terms = stage_copy(c.bwd.terms, c.bwd.n_terms);
      }
      if constexpr ((STAGE & 2) != 0)
      {
        T.lu_tasks = reinterpret_cast<const uint2*>(
            stage_copy(reinterpret_cast<const uint32_t*>(c.lu.tasks), 2 * (static_cast<size_t>(c.lu.n_tasks) + 1)));
        T.lu_terms = stage_copy(c.lu.terms, c.lu.n_terms);
      }
      coop_sync<NT>();
    }
    while (true)
    {
      if (lane == 0)
        s_idx = atomicAdd(next, 1);
      coop_sync<NT>();
      const int64_t idx = s_idx;
      coop_sync<NT>();
      if (idx >= count)
        return;
      const int64_t cell = list ? list[idx] : idx;
      const size_t g = cell / kGroup;
      const int cl = cell % kGroup;
      {
        Lane ys = lane_view(y_state, g, n, cl);
        Lane ks = lane_view(k_state, g, m.n_reactions, cl);
        for (int i = lane; i < n; i += NT)
          L.y[i] = ys[i];
        if (cond_state)
        {
          // UpdateStateParameters fused into the launch
          Lane cd = lane_view(const_cast<double*>(cond_state), g, 3, cl);
          Lane cu = lane_view(const_cast<double*>(custom_state), g, m.n_custom, cl);
          const double T = cd[0], P = cd[1], rho = cd[2];
          for (int r = lane; r < m.n_reactions; r += NT)
          {
            const double kr = rate_constant(m, r, T, P, rho, cu);
            L.k[r] = kr;
            ks[r] = kr;
          }
        }
        else
        {
          for (int r = lane; r < m.n_reactions; r += NT)
            L.k[r] = ks[r];
        }
      }
      CellControl cc;
      if (resume)
        cc = resume[cell];
      else
        control_init(p, cc);
      // recorded Solve: the cell's checkpoints are appended to its tape
      int64_t n_rec = RECORD ? tape.counts[cell] : 0;
      double* const rec_base =
          RECORD ? tape.records + static_cast<size_t>(cell) * tape.max_records * (static_cast<size_t>(n) + 1) : nullptr;
      auto record = [&](double tag) {
        if (n_rec < tape.max_records)
        {
          double* rec = rec_base + static_cast<size_t>(n_rec) * (static_cast<size_t>(n) + 1);
          if (lane == 0)
            rec[0] = tag;
          for (int i = lane; i < n; i += NT)
            rec[1 + i] = L.y[i];
        }
        ++n_rec;
      };
      coop_sync<NT>();
      while (true)
      {
        bool fresh = false;
        if (cc.flags & kNewStep)
        {
          if (!begin_step(p, cc))
            break;
          if constexpr (COMPACT)
          {
            coop_forcing<NT>(m, c, L, L.y, L.K);
            if (lane < n)
              f0_r = L.K[lane];
          }
          else
          {
            coop_forcing<NT>(m, c, L, L.y, L.f0);
            coop_jacobian<NT>(m, c, L);
          }
          cc.stats[MICM_HIP_STAT_FUNCTION_CALLS] += 1;
          cc.stats[MICM_HIP_STAT_JACOBIAN_UPDATES] += 1;
          cc.last_alpha = 0.0;
          cc.flags &= ~kNewStep;
          fresh = true;
        }
        const double alpha = next_alpha(p, cc);
        if constexpr (COMPACT)
        {
          // -J(y) again (y and k are those of the step), its diagonal replaced by the
          // accumulated one: the same values as keeping -J and adding alpha in place
          // (and the zero one past the factor that the LU records' padded slots read)
          for (int q = lane; q <= c.nnz_alu; q += NT)
            L.alu[q] = 0.0;
          coop_sync<NT>();
          coop_jacobian<NT, true>(m, c, L);
          if (lane < n)
          {
            const int d = c.alu_diag[lane];
            if (fresh)
              jdiag_r = L.alu[d];
            jdiag_r += alpha * 1.0;
            L.alu[d] = jdiag_r;
          }
          coop_sync<NT>();
          coop_levels_piped<NT, true>(c.lu, c.lu.terms, L.alu, L.alu, L.alu);
        }
        else
        {
          for (int i = lane; i < n; i += NT)
            L.jac[m.jac_diag[i]] += alpha * 1.0;
          coop_sync<NT>();
          coop_decompose<NT, (STAGE & 2) == 0>(m, c, L, T);
        }
        cc.stats[MICM_HIP_STAT_DECOMPOSITIONS] += 1;
        for (int s = 0; s < p.stages; ++s)
        {
          const int sc = s * (s - 1) / 2;
          double* Ks = L.K + static_cast<size_t>(s) * n;
          if (s == 0)
          {
            for (int i = lane; i < n; i += NT)
              Ks[i] = COMPACT ? f0_r : L.f0[i];
          }
          else if (p.new_f[s])
          {
            // compact: the forcing's input goes into its output vector (the forcing
            // reads its input before it writes)
            double* yin = COMPACT ? Ks : L.ynew;
            for (int i = lane; i < n; i += NT)
            {
              double v = L.y[i];
              for (int j = 0; j < s; ++j)
                v += p.a[sc + j] * L.K[static_cast<size_t>(j) * n + i];
              yin[i] = v;
            }
            coop_sync<NT>();
            coop_forcing<NT>(m, c, L, yin, Ks);
            cc.stats[MICM_HIP_STAT_FUNCTION_CALLS] += 1;
          }
          coop_sync<NT>();
          if (s + 1 < p.stages && !p.new_f[s + 1])
            for (int i = lane; i < n; i += NT)
              Ks[static_cast<size_t>(n) + i] = Ks[i];
          if (s > 0)
          {
            double coef[6];
            for (int j = 0; j < s; ++j)
              coef[j] = p.c[sc + j] / cc.H;
            for (int i = lane; i < n; i += NT)
            {
              double v = Ks[i];
              for (int j = 0; j < s; ++j)
                v += coef[j] * L.K[static_cast<size_t>(j) * n + i];
              Ks[i] = v;
            }
          }
          coop_sync<NT>();
          coop_solve<NT, COMPACT || (STAGE & 1) == 0>(c, L, T, Ks);
          cc.stats[MICM_HIP_STAT_SOLVES] += 1;
        }
        for (int i = lane; i < n; i += NT)
        {
          double v = L.y[i], e = 0.0;
          for (int s = 0; s < p.stages; ++s)
            v += p.m[s] * L.K[static_cast<size_t>(s) * n + i];
          for (int s = 0; s < p.stages; ++s)
            e += p.e[s] * L.K[static_cast<size_t>(s) * n + i];
          if constexpr (COMPACT)
          {
            ynew_r = v;
          }
          else
          {
            L.ynew[i] = v;
            L.yerr[i] = e;
          }
          const double a = fabs(L.y[i]), b = fabs(v);
          const double q = e / (atol[i] + p.rtol * (a > b ? a : b));
          L.scratch[i] = q * q;
        }
        coop_sync<NT>();
        if (lane == 0)
        {
          double acc = 0.0;
          for (int i = 0; i < n; ++i)
            acc += L.scratch[i];
          s_error = smax(sqrt(acc / static_cast<double>(n > 0 ? n : 1)), 1.0e-10);
        }
        coop_sync<NT>();
        const double error = s_error;
        const double h_used = cc.H;
        if (control_update(p, cc, error))
        {
          if constexpr (RECORD)
            record(h_used);
          for (int i = lane; i < n; i += NT)
            L.y[i] = COMPACT ? ynew_r : L.ynew[i];
        }
        coop_sync<NT>();
        if (cc.status != MICM_HIP_STATE_RUNNING)
          break;
      }
      if constexpr (RECORD)
      {
        record(kClampTag);
        if (lane == 0)
          tape.counts[cell] = n_rec;
      }
      {
        Lane ys = lane_view(y_state, g, n, cl);
        for (int i = lane; i < n; i += NT)
          ys[i] = smax(L.y[i], 0.0);
      }
      if (lane == 0)
      {
        status_out[cell] = cc.status;
        for (int i = 0; i < 7; ++i)
          stats_out[cell * 7 + i] = cc.stats[i];
        final_time_out[cell] = cc.t;
      }
      coop_sync<NT>();
    }
  }
}  // namespace

// ====================================================================== C-ABI

// Scratch of micm_hip_solve_noalloc: the cooperative launch's queue counter and the compact
// layout's per-workgroup rate-constant rows.  Each GPU holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress reports for the
// current device (noalloc_scratch).
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
constexpr size_t kNoAllocCounterBytes = 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace
{
  constexpr int kMaxDevices = 64;
  std::mutex g_scratch_mutex;
  void* g_scratch_address[kMaxDevices] = {};

  // The current device's scratch address, looked up once per device.
  int noalloc_scratch(void** out)
  {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
      return fail(MICM_HIP_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[dev])
    {
      void* p = nullptr;
      HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      if (!p)
        return fail(MICM_HIP_ERR_DEVICE, "no address for the scratch array");
      g_scratch_address[dev] = p;
    }
    *out = g_scratch_address[dev];
    return MICM_HIP_OK;
  }
}  // namespace

struct micm_hip_mechanism
{
  MechanismTables t;
  DevMech dev{};
  DevCoop coop{};
  DevLevels lu_blocked[kNumLuBlockings] = {};  // the LU schedule of each blocking (kLuBlockings)
  // lu_blocked[b] with levels capped at kLuCaps[c] tasks, at b * kNumLuCaps + c (offsets
  // null where no level of lu_blocked[b] exceeds the cap)
  DevLevels lu_capped[kNumLuBlockings * kNumLuCaps] = {};
  DevAdjoint adj{};
  void* blob = nullptr;  // one device allocation for every table
  int device = 0;
};

struct micm_hip_state
{
  const micm_hip_mechanism* mech = nullptr;
  int64_t n_cells = 0, n_groups = 0;
  hipStream_t stream = nullptr;
  double rtol = 1e-6;
  // device arrays
  double *y = nullptr, *cond = nullptr, *custom = nullptr, *k = nullptr, *jac = nullptr, *lower = nullptr,
         *upper = nullptr, *atol = nullptr;
  double *f0 = nullptr, *ynew = nullptr, *yerr = nullptr, *K = nullptr;
  CellControl* ctl = nullptr;
  int32_t* status = nullptr;
  int64_t* stats = nullptr;
  double* final_time = nullptr;
  unsigned int* active = nullptr;
  double* lockstep_error = nullptr;
  int* queue_next = nullptr;
  unsigned int* host_active = nullptr;  // pinned
  // persistent scratch
  int64_t n_slots = 0;
  double* slot_blob = nullptr;
  int32_t* order = nullptr;  // persistent processing order
  bool has_order = false;
  // hand-off to the cooperative form
  int64_t handoff_steps = -1;
  int32_t* handoff_list = nullptr;
  int* handoff_count = nullptr;
  int* coop_next = nullptr;
  int* order_bins = nullptr;  // kOrderBins + 1: the processing order's counting sort
  double* coop_k = nullptr;  // rate-constant rows of the compact cooperative layout
  int64_t coop_k_rows = 0;
  // cooperative launch shape (-1 / 0: chosen from the mechanism)
  int coop_stage = -1;
  int coop_threads = 0;
  // shortest LU segment of the cooperative form's schedule (0: kLuBlockingDefault; -1:
  // whole sums)
  int lu_blocking = 0;
  // the tape a recorded Solve appends to (records == nullptr otherwise)
  Tape recording{ nullptr, nullptr, 0 };
  // vector length of the dense matrix policy whose summation order the lockstep form's
  // error norm follows (1: Matrix<Real>)
  int64_t lockstep_group = 1;
  // scratch of the launch being enqueued by micm_hip_solve_noalloc (nullptr otherwise)
  int* ext_next = nullptr;
  double* ext_k = nullptr;
  int64_t ext_k_rows = 0;
};

struct micm_hip_tape
{
  const micm_hip_mechanism* mech = nullptr;
  int64_t n_cells = 0, max_records = 0;
  double* records = nullptr;
  int64_t* counts = nullptr;
};

struct micm_hip_packed
{
  std::vector<micm_hip_state*> states;
  int32_t form = MICM_HIP_SOLVE_COOPERATIVE;
  hipStream_t origin = nullptr;      // capture and launch stream
  hipStream_t own_stream = nullptr;  // the stream created with the object
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // what the captured graph was built for
  bool captured = false, replayed = false;
  double dt = 0.0;
  micm_hip_rosenbrock_parameters params{};
  std::vector<int64_t> shapes;  // per state: launch shapes, rtol and scratch buffers the graph bakes in
};

namespace
{
  template<class T>
  int upload(void* blob, size_t& off, const std::vector<T>& v, const T*& dst, std::vector<char>& staging)
  {
    off = (off + 255) & ~size_t(255);
    const size_t bytes = v.size() * sizeof(T);
    if (staging.size() < off + bytes)
      staging.resize(off + bytes);
    if (bytes)
      std::memcpy(staging.data() + off, v.data(), bytes);
    dst = reinterpret_cast<const T*>(static_cast<char*>(blob) + off);
    off += bytes;
    return 0;
  }

  int upload_pairs(void* blob, size_t& off, const std::vector<int32_t>& v, const int2*& dst, std::vector<char>& staging)
  {
    const int32_t* p = nullptr;
    upload(blob, off, v, p, staging);
    dst = reinterpret_cast<const int2*>(p);
    return 0;
  }
  void upload_levels(void* blob, size_t& off, const LevelSchedule& l, DevLevels& d, std::vector<char>& staging)
  {
    upload(blob, off, l.offsets, d.offsets, staging);
    const uint32_t* tasks = nullptr;
    upload(blob, off, l.tasks, tasks, staging);
    d.tasks = reinterpret_cast<const uint2*>(tasks);
    upload(blob, off, l.terms, d.terms, staging);
    const uint32_t* records = nullptr;
    upload(blob, off, l.records, records, staging);
    d.records = reinterpret_cast<const uint4*>(records);
    d.n_levels = l.n_levels();
    d.n_tasks = l.tasks.empty() ? 0 : l.n_tasks();
    d.n_terms = static_cast<int32_t>(l.terms.size());
    d.max_level_terms = l.max_level_terms;
  }


  template<class T>
  size_t padded_bytes(const std::vector<T>& v)
  {
    return ((v.size() * sizeof(T) + 255) & ~size_t(255)) + 256;
  }

  size_t level_bytes(const LevelSchedule& l)
  {
    return padded_bytes(l.offsets) + padded_bytes(l.tasks) + padded_bytes(l.terms) + padded_bytes(l.records);
  }

  size_t blocked_level_bytes(const MechanismTables& t)
  {
    size_t bytes = 0;
    for (const LevelSchedule& l : t.lv_lu_blocked)
      bytes += level_bytes(l);
    for (const LevelSchedule& l : t.lv_lu_capped)
      if (l.n_levels() > 0)
        bytes += level_bytes(l);
    return bytes;
  }

  int check_mech(const micm_hip_mechanism* m)
  {
    if (!m)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null mechanism");
    return MICM_HIP_OK;
  }

  int ensure_uploaded(const micm_hip_mechanism* m);

  // A mechanism whose tables are on the device.
  int check_device_mech(const micm_hip_mechanism* m)
  {
    if (int rc = check_mech(m))
      return rc;
    return ensure_uploaded(m);
  }

  DevParams make_params(const micm_hip_rosenbrock_parameters& in, double dt, double rtol)
  {
    DevParams p{};
    p.stages = static_cast<int32_t>(in.stages);
    p.max_steps = in.max_number_of_steps;
    p.round_off = in.round_off;
    p.factor_min = in.factor_min;
    p.factor_max = in.factor_max;
    p.reject_decrease = in.rejection_factor_decrease;
    p.safety = in.safety_factor;
    // rosenbrock.inl:33-35
    p.h_min = in.h_min == 0.0 ? 1.0e-15 * dt : in.h_min;
    p.h_max = in.h_max == 0.0 ? dt : std::min(dt, in.h_max);
    p.h_start = in.h_start == 0.0 ? 1.0e-6 * dt : std::min(p.h_max, in.h_start);
    p.inv_elo = 1 / in.estimator_of_local_order;
    p.gamma0 = in.gamma[0];
    p.dt = dt;
    p.rtol = rtol;
    p.error_group = 1;
    for (int i = 0; i < 6; ++i)
    {
      p.new_f[i] = in.new_function_evaluation[i];
      p.m[i] = in.m[i];
      p.e[i] = in.e[i];
    }
    for (int i = 0; i < 15; ++i)
    {
      p.a[i] = in.a[i];
      p.c[i] = in.c[i];
    }
    return p;
  }

  Work state_work(micm_hip_state* s)
  {
    const auto& m = s->mech->dev;
    Work w{};
    w.y = s->y;
    w.k = s->k;
    w.jac = s->jac;
    w.lower = s->lower;
    w.upper = s->upper;
    w.f0 = s->f0;
    w.ynew = s->ynew;
    w.yerr = s->yerr;
    w.K = s->K;
    w.k_stride = static_cast<size_t>(s->n_groups) * m.n_species * kGroup;
    return w;
  }

  void host_to_grouped(const double* host, int64_t n_cells, int64_t n_vals, std::vector<double>& out)
  {
    const int64_t ng = n_groups_of(n_cells);
    out.assign(static_cast<size_t>(ng) * n_vals * kGroup, 0.0);
    for (int64_t c = 0; c < n_cells; ++c)
      for (int64_t v = 0; v < n_vals; ++v)
        out[(static_cast<size_t>(c / kGroup) * n_vals + v) * kGroup + c % kGroup] = host[c * n_vals + v];
  }

  void grouped_to_host(const std::vector<double>& in, int64_t n_cells, int64_t n_vals, double* host)
  {
    for (int64_t c = 0; c < n_cells; ++c)
      for (int64_t v = 0; v < n_vals; ++v)
        host[c * n_vals + v] = in[(static_cast<size_t>(c / kGroup) * n_vals + v) * kGroup + c % kGroup];
  }

  size_t grouped_doubles(const micm_hip_state* s, int64_t n_vals)
  {
    return static_cast<size_t>(s->n_groups) * static_cast<size_t>(n_vals) * kGroup;
  }

  int set_grouped(micm_hip_state* s, double* dst, const double* host, int64_t n_vals)
  {
    if (!s || (!host && n_vals))
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null argument");
    if (n_vals == 0)
      return MICM_HIP_OK;
    std::vector<double> buf;
    host_to_grouped(host, s->n_cells, n_vals, buf);
    HIP_TRY(hipMemcpyAsync(dst, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MICM_HIP_OK;
  }

  int get_grouped(const micm_hip_state* s, const double* src, double* host, int64_t n_vals)
  {
    if (!s || (!host && n_vals))
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null argument");
    if (n_vals == 0)
      return MICM_HIP_OK;
    std::vector<double> buf(static_cast<size_t>(s->n_groups) * n_vals * kGroup);
    HIP_TRY(hipMemcpyAsync(buf.data(), src, buf.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    grouped_to_host(buf, s->n_cells, n_vals, host);
    return MICM_HIP_OK;
  }

  int ensure_slots(micm_hip_state* s, int64_t slots, int stages)
  {
    (void)stages;
    if (s->n_slots >= slots)
      return MICM_HIP_OK;
    if (s->slot_blob)
      HIP_TRY(hipFree(s->slot_blob));
    s->slot_blob = nullptr;
    const auto& m = s->mech->dev;
    const size_t per = static_cast<size_t>(m.n_reactions + m.nnz_jac + m.nnz_lower + m.nnz_upper + 10 * m.n_species);
    HIP_TRY(hipMalloc(&s->slot_blob, per * kGroup * sizeof(double) * slots));
    s->n_slots = slots;
    return MICM_HIP_OK;
  }

  Work slot_work(micm_hip_state* s)
  {
    const auto& m = s->mech->dev;
    const size_t S = static_cast<size_t>(s->n_slots) * kGroup;
    double* p = s->slot_blob;
    Work w{};
    auto take = [&](int64_t n) {
      double* r = p;
      p += S * n;
      return r;
    };
    w.y = take(m.n_species);
    w.k = take(m.n_reactions);
    w.jac = take(m.nnz_jac);
    w.lower = take(m.nnz_lower);
    w.upper = take(m.nnz_upper);
    w.f0 = take(m.n_species);
    w.ynew = take(m.n_species);
    w.yerr = take(m.n_species);
    w.K = take(6 * static_cast<int64_t>(m.n_species));
    w.k_stride = S * m.n_species;
    return w;
  }

  // Copies every table of a compiled mechanism into one device allocation.  Runs on
  // the first call that needs the device, so a mechanism can be compiled and
  // inspected on a host without a GPU.
  int ensure_uploaded(const micm_hip_mechanism* cm)
  {
    auto* m = const_cast<micm_hip_mechanism*>(cm);
    if (m->blob)
      return MICM_HIP_OK;
    const MechanismTables& t = m->t;
    HIP_TRY(hipInit(0));
    HIP_TRY(hipGetDevice(&m->device));
    size_t bytes = 0;
    bytes += padded_bytes(t.rate_type) + padded_bytes(t.rate_params) + padded_bytes(t.custom_index) +
             padded_bytes(t.multiplier_offsets) + padded_bytes(t.multipliers) + padded_bytes(t.reactant_offsets) +
             padded_bytes(t.reactant_ids) + padded_bytes(t.product_offsets) + padded_bytes(t.product_ids) +
             padded_bytes(t.yields) + padded_bytes(t.term_process) + padded_bytes(t.term_dep_offsets) +
             padded_bytes(t.term_dep_ids) + padded_bytes(t.term_flat_offsets) + padded_bytes(t.term_flat) +
             padded_bytes(t.jac_diag) + padded_bytes(t.lu_out) + padded_bytes(t.lu_a) + padded_bytes(t.lu_div) +
             padded_bytes(t.lu_term_offsets) + padded_bytes(t.lu_term_pairs) +
             padded_bytes(t.lower_diag) + padded_bytes(t.upper_diag) + padded_bytes(t.fwd_offsets) +
             padded_bytes(t.fwd_pairs) + padded_bytes(t.bwd_offsets) + padded_bytes(t.bwd_pairs) +
             padded_bytes(t.jac_to_alu) + padded_bytes(t.alu_diag) + padded_bytes(t.alu_lower) +
             padded_bytes(t.alu_upper) + level_bytes(t.lv_lu) + blocked_level_bytes(t) + level_bytes(t.lv_fwd) +
             padded_bytes(t.fwd_sweep.lanes) + level_bytes(t.lv_bwd) + padded_bytes(t.bwd_chain.headers) +
             padded_bytes(t.bwd_chain.lane_terms) +
             padded_bytes(t.rxn_pack) + padded_bytes(t.term_pack) +
             padded_bytes(t.spec_offsets) + padded_bytes(t.spec_rxn) +
             padded_bytes(t.spec_coef) + padded_bytes(t.jel_offsets) + padded_bytes(t.jel_term) +
             padded_bytes(t.jel_coef) + padded_bytes(t.jel_pack) + level_bytes(t.lv_ut) + level_bytes(t.lv_lt) +
             padded_bytes(t.term_indep) + padded_bytes(t.rxn_term_offsets) + padded_bytes(t.rxn_terms) +
             padded_bytes(t.indep_term_offsets) + padded_bytes(t.indep_terms) + padded_bytes(t.hes_offsets) +
             padded_bytes(t.hes_term) + padded_bytes(t.hes_slot) + 256;
    HIP_TRY(hipMalloc(&m->blob, bytes));
    std::vector<char> staging;
    size_t off = 0;
    DevMech& d = m->dev;
    d.n_species = static_cast<int32_t>(t.n_species);
    d.n_reactions = static_cast<int32_t>(t.n_reactions);
    d.n_custom = static_cast<int32_t>(t.n_custom);
    d.nnz_jac = static_cast<int32_t>(t.nnz_jacobian());
    d.nnz_lower = static_cast<int32_t>(t.nnz_lower());
    d.nnz_upper = static_cast<int32_t>(t.nnz_upper());
    d.n_terms = static_cast<int32_t>(t.term_process.size());
    d.n_lu_elements = static_cast<int32_t>(t.lu_out.size());
    upload(m->blob, off, t.rate_type, d.rate_type, staging);
    upload(m->blob, off, t.rate_params, d.rate_params, staging);
    upload(m->blob, off, t.custom_index, d.custom_index, staging);
    upload(m->blob, off, t.multiplier_offsets, d.multiplier_offsets, staging);
    upload(m->blob, off, t.multipliers, d.multipliers, staging);
    upload(m->blob, off, t.reactant_offsets, d.reactant_offsets, staging);
    upload(m->blob, off, t.reactant_ids, d.reactant_ids, staging);
    upload(m->blob, off, t.product_offsets, d.product_offsets, staging);
    upload(m->blob, off, t.product_ids, d.product_ids, staging);
    upload(m->blob, off, t.yields, d.yields, staging);
    upload(m->blob, off, t.term_process, d.term_process, staging);
    upload(m->blob, off, t.term_dep_offsets, d.term_dep_offsets, staging);
    upload(m->blob, off, t.term_dep_ids, d.term_dep_ids, staging);
    upload(m->blob, off, t.term_flat_offsets, d.term_flat_offsets, staging);
    upload(m->blob, off, t.term_flat, d.term_flat, staging);
    upload(m->blob, off, t.jac_diag, d.jac_diag, staging);
    upload(m->blob, off, t.lu_out, d.lu_out, staging);
    upload(m->blob, off, t.lu_a, d.lu_a, staging);
    upload(m->blob, off, t.lu_div, d.lu_div, staging);
    upload(m->blob, off, t.lu_term_offsets, d.lu_term_offsets, staging);
    upload_pairs(m->blob, off, t.lu_term_pairs, d.lu_terms, staging);
    upload(m->blob, off, t.lower_diag, d.lower_diag, staging);
    upload(m->blob, off, t.upper_diag, d.upper_diag, staging);
    upload(m->blob, off, t.fwd_offsets, d.fwd_offsets, staging);
    upload_pairs(m->blob, off, t.fwd_pairs, d.fwd_terms, staging);
    upload(m->blob, off, t.bwd_offsets, d.bwd_offsets, staging);
    upload_pairs(m->blob, off, t.bwd_pairs, d.bwd_terms, staging);
    DevCoop& c = m->coop;
    c.nnz_alu = static_cast<int32_t>(t.nnz_alu());
    c.n_terms = static_cast<int32_t>(t.term_process.size());
    upload(m->blob, off, t.jac_to_alu, c.jac_to_alu, staging);
    upload(m->blob, off, t.alu_diag, c.alu_diag, staging);
    upload(m->blob, off, t.alu_lower, c.alu_lower, staging);
    upload(m->blob, off, t.alu_upper, c.alu_upper, staging);
    upload_levels(m->blob, off, t.lv_lu, c.lu, staging);
    for (size_t b = 0; b < t.lv_lu_blocked.size() && b < static_cast<size_t>(kNumLuBlockings); ++b)
      upload_levels(m->blob, off, t.lv_lu_blocked[b], m->lu_blocked[b], staging);
    for (size_t k = 0; k < t.lv_lu_capped.size() && k < static_cast<size_t>(kNumLuBlockings * kNumLuCaps); ++k)
      if (t.lv_lu_capped[k].n_levels() > 0)
        upload_levels(m->blob, off, t.lv_lu_capped[k], m->lu_capped[k], staging);
    upload_levels(m->blob, off, t.lv_fwd, c.fwd, staging);
    c.fwd_sweep = DevSweep{ nullptr, -1, 0 };
    if (t.fwd_sweep.first_level >= 0)
    {
      const uint32_t* lanes = nullptr;
      upload(m->blob, off, t.fwd_sweep.lanes, lanes, staging);
      c.fwd_sweep = DevSweep{ reinterpret_cast<const uint4*>(lanes), t.fwd_sweep.first_level, t.fwd_sweep.n_cols };
    }
    upload_levels(m->blob, off, t.lv_bwd, c.bwd, staging);
    c.bwd_chain = DevChain{ nullptr, nullptr, 0, 0 };
    if (t.bwd_chain.levels > 0)
    {
      upload(m->blob, off, t.bwd_chain.headers, c.bwd_chain.headers, staging);
      upload(m->blob, off, t.bwd_chain.lane_terms, c.bwd_chain.lane_terms, staging);
      c.bwd_chain.levels = t.bwd_chain.levels;
      c.bwd_chain.n_tasks = static_cast<int32_t>(t.bwd_chain.lane_terms.size() / kChainLanes);
    }
    upload(m->blob, off, t.rxn_pack, c.rxn_pack, staging);
    upload(m->blob, off, t.term_pack, c.term_pack, staging);
    upload(m->blob, off, t.spec_offsets, c.spec_offsets, staging);
    upload(m->blob, off, t.spec_rxn, c.spec_rxn, staging);
    upload(m->blob, off, t.spec_coef, c.spec_coef, staging);
    upload(m->blob, off, t.jel_offsets, c.jel_offsets, staging);
    upload(m->blob, off, t.jel_term, c.jel_term, staging);
    upload(m->blob, off, t.jel_coef, c.jel_coef, staging);
    upload(m->blob, off, t.jel_pack, c.jel_pack, staging);
    DevAdjoint& a = m->adj;
    upload_levels(m->blob, off, t.lv_ut, a.ut, staging);
    upload_levels(m->blob, off, t.lv_lt, a.lt, staging);
    upload(m->blob, off, t.term_indep, a.term_indep, staging);
    upload(m->blob, off, t.rxn_term_offsets, a.rxn_term_offsets, staging);
    upload(m->blob, off, t.rxn_terms, a.rxn_terms, staging);
    upload(m->blob, off, t.indep_term_offsets, a.indep_term_offsets, staging);
    upload(m->blob, off, t.indep_terms, a.indep_terms, staging);
    upload(m->blob, off, t.hes_offsets, a.hes_offsets, staging);
    upload(m->blob, off, t.hes_term, a.hes_term, staging);
    upload(m->blob, off, t.hes_slot, a.hes_slot, staging);
    if (off > bytes)
      return fail(MICM_HIP_ERR_DEVICE, "internal table sizing error");
    HIP_TRY(hipMemcpy(m->blob, staging.data(), off, hipMemcpyHostToDevice));
    return MICM_HIP_OK;
  }

  // Launch plan of the cooperative kernel: dynamic LDS, which tables are staged in
  // LDS, threads per cell.  lds == 0: the mechanism does not fit.
  struct CoopPlan
  {
    size_t lds = 0;
    int stage = 0;
    int threads = kWave;
    int p_capacity = 0;     // doubles of the per-level product buffer
    bool compact = false;   // coop_layout_compact
    // compact, 256 threads: the build's register budget as waves per SIMD (3: 168 registers,
    // 2: 256, 1: unbounded); 0: the NT's default build
    int min_waves = 0;
  };

  // Workgroups of the cooperative kernel that fit one CU (registers and LDS).
  template<int NT>
  int coop_occupancy_nt(int stage, bool compact, size_t lds, int min_waves = 0)
  {
    int per_cu = 0;
    hipError_t e = hipSuccess;
    if constexpr (NT == 256)
    {
      if (compact && (min_waves == 1 || min_waves == 2))
      {
        e = min_waves == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_coop<NT, 0, true, 1>, NT, lds)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_coop<NT, 0, true, 2>, NT, lds);
        return e == hipSuccess ? per_cu : 0;
      }
    }
    if (compact)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_coop<NT, 0, true>, NT, lds);
    else if (stage == 1)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_coop<NT, 1>, NT, lds);
    else if (stage == 2)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_coop<NT, 2>, NT, lds);
    else if (stage == 3)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_coop<NT, 3>, NT, lds);
    else
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_coop<NT, 0>, NT, lds);
    return e == hipSuccess ? per_cu : 0;
  }

  int coop_occupancy(int threads, int stage, bool compact, size_t lds, int min_waves = 0)
  {
    switch (threads)
    {
      case 128: return coop_occupancy_nt<128>(stage, compact, lds);
      case 256: return coop_occupancy_nt<256>(stage, compact, lds, min_waves);
      case 512: return coop_occupancy_nt<512>(stage, compact, lds);
      case 1024: return coop_occupancy_nt<1024>(stage, compact, lds);
      default: return coop_occupancy_nt<kWave>(stage, compact, lds);
    }
  }

  size_t lds_limit()
  {
    int dev = 0, max_lds = 0, optin = 0;
    if (hipGetDevice(&dev) != hipSuccess)
      return 0;
    (void)hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    (void)hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev);
    return static_cast<size_t>(std::max(max_lds, optin)) - 64;  // static __shared__ of the kernel
  }

  // Cells per workgroup of the tiled LU / substitution kernels: the largest power of
  // two up to 16 whose tile of `doubles_per_cell` leaves room for two workgroups per CU
  // (two independent workgroups hide each other's level latency: 210/560 LU 1.4x,
  // 60/200 2.1x against one workgroup of twice the cells), else one cell if it fits;
  // 0 when the level schedules are unavailable or one cell does not fit.
  int tile_cells(const micm_hip_mechanism* m, size_t doubles_per_cell)
  {
    if (!m->t.coop_levels_ok)
      return 0;
    const size_t limit = lds_limit();
    for (int tc = 16; tc >= 1; tc /= 2)
      if (2 * doubles_per_cell * tc * sizeof(double) <= limit)
        return tc;
    return doubles_per_cell * sizeof(double) <= limit ? 1 : 0;
  }

  // Cells per workgroup of the tiled forcing / Jacobian kernels: the largest power of two
  // from 16 down to 2 whose tile of y and k leaves room for three workgroups per CU
  // (two for the smallest tile); 0 when the gather tables are unavailable or nothing fits.
  constexpr int kGatherThreads = 512;

  size_t gather_tile_bytes(const micm_hip_mechanism* m, int tc)
  {
    return (static_cast<size_t>(m->dev.n_species) + static_cast<size_t>(m->dev.n_reactions)) * tc * sizeof(double);
  }

  int gather_tile_cells(const micm_hip_mechanism* m)
  {
    if (!m->t.coop_levels_ok)
      return 0;
    const size_t limit = lds_limit();
    for (int tc = 16; tc >= 4; tc /= 2)
      if (3 * gather_tile_bytes(m, tc) <= limit)
        return tc;
    return 2 * gather_tile_bytes(m, 2) <= limit ? 2 : 0;
  }

  inline unsigned tiled_blocks(int tc, int64_t n_cells)
  {
    return static_cast<unsigned>((n_groups_of(n_cells) + kXcds - 1) / kXcds * kXcds * (kGroup / tc));
  }

#define LAUNCH_GATHER(kernel, tc, lds, stream, ...)                                                                  \
  do                                                                                                                \
  {                                                                                                                 \
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel<tc, kGatherThreads>),                          \
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));                \
    hipLaunchKernelGGL((kernel<tc, kGatherThreads>), dim3(tiled_blocks(tc, n_cells)), dim3(kGatherThreads), lds,    \
                       stream, __VA_ARGS__);                                                                        \
  } while (0)

#define LAUNCH_GATHER_Z(kernel, tc, zero, lds, stream, ...)                                                         \
  do                                                                                                                \
  {                                                                                                                 \
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel<tc, kGatherThreads, zero>),                    \
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));                \
    hipLaunchKernelGGL((kernel<tc, kGatherThreads, zero>), dim3(tiled_blocks(tc, n_cells)), dim3(kGatherThreads),   \
                       lds, stream, __VA_ARGS__);                                                                   \
  } while (0)

#define LAUNCH_TILED(kernel, blocks, lds, stream, ...)                                                               \
  do                                                                                                                \
  {                                                                                                                 \
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, \
                                static_cast<int>(lds)));                                                            \
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kTileThreads), lds, stream, __VA_ARGS__);                        \
  } while (0)

  // Index into kLuBlockings of an LU blocking as micm_hip_state_set_lu_blocking takes it
  // (0: the default, -1: whole sums), or -1 for a blocking that is not built.
  int lu_blocking_index(int blocking)
  {
    const int want = blocking == 0 ? kLuBlockingDefault : blocking == -1 ? 0 : blocking;
    if (want < 0)
      return -1;
    for (int b = 0; b < kNumLuBlockings; ++b)
      if (kLuBlockings[b] == want)
        return b;
    return -1;
  }

  // The cooperative tables with the LU schedule of `blocking` for a workgroup of `threads`:
  // its levels capped at `threads` tasks where that schedule exists.
  DevCoop coop_with_blocking(const micm_hip_mechanism* m, int blocking, int threads)
  {
    DevCoop c = m->coop;
    const int b = lu_blocking_index(blocking);
    if (!m->t.coop_levels_ok || b < 0)
      return c;
    if (m->lu_blocked[b].offsets)
      c.lu = m->lu_blocked[b];
    for (int k = 0; k < kNumLuCaps; ++k)
      if (kLuCaps[k] == threads && m->lu_capped[b * kNumLuCaps + k].offsets)
        c.lu = m->lu_capped[b * kNumLuCaps + k];
    return c;
  }

  // Threads of a cooperative workgroup as plan_coop takes them: 128 ... 1024, else one
  // wavefront.
  int coop_threads_of(int threads_request)
  {
    return (threads_request == 128 || threads_request == 256 || threads_request == 512 || threads_request == 1024)
               ? threads_request
               : kWave;
  }

  CoopPlan plan_coop_full(const micm_hip_mechanism* m, int stages, int stage_request, int threads_request,
                          int lu_blocking);

  // stage_request < 0: stage the substitution tables when they fit and the LU tables
  // when that costs no resident cell per CU, or take the compact layout when that fits
  // more cells per CU; 4: the compact layout; threads_request 0: one wavefront.
  CoopPlan plan_coop(const micm_hip_mechanism* m, int stages, int stage_request = -1, int threads_request = 0,
                     int lu_blocking = 0)
  {
    // 4: the compact layout, its register budget chosen by occupancy; 5 / 6 / 7: the compact
    // layout built for three / two waves per SIMD / without a register bound (256 threads)
    const bool compact_request = stage_request >= 4;
    CoopPlan plan = plan_coop_full(m, stages, compact_request ? -1 : stage_request, threads_request, lu_blocking);
    if (!m->t.coop_levels_ok || (stage_request >= 0 && !compact_request))
      return plan;
    // the compact layout keeps one species per thread in registers and publishes the
    // rate constants through global memory, so it needs a workgroup barrier
    const int threads = coop_threads_of(threads_request);
    const DevCoop c = coop_with_blocking(m, lu_blocking, threads);
    if (threads <= kWave || threads < m->dev.n_species)
      return plan;
    const size_t compact = coop_compact_lds_doubles(c.nnz_alu, m->dev.n_species, stages, m->dev.n_reactions,
                                                    c.n_terms, c.bwd.max_level_terms) *
                           sizeof(double);
    if (compact > lds_limit())
      return plan;
    // the 256-thread compact build in three register budgets: the one that fits the most cells
    // per CU, the largest budget among equals (fewer spills)
    int occ[4] = { 0, 0, 0, 0 };
    int best_waves = 0;
    for (int w = 1; w <= 3; ++w)
    {
      if (threads != 256 && w != 3)
        continue;
      occ[w] = coop_occupancy(threads, 0, true, compact, w);
      if (best_waves == 0 || occ[w] > occ[best_waves])
        best_waves = w;
    }
    const int occ_compact = occ[best_waves];
    const int occ_full = plan.lds > 0 ? coop_occupancy(threads, plan.stage, false, plan.lds) : 0;
    if (compact_request ? occ_compact > 0 : occ_compact > occ_full)
    {
      plan.lds = compact;
      plan.stage = 0;
      plan.threads = threads;
      plan.p_capacity = c.bwd.max_level_terms;
      plan.compact = true;
      plan.min_waves = threads == 256 ? best_waves : 0;
      // 5 / 6 / 7: three / two waves per SIMD / no register bound, where that build fits
      const int forced = stage_request == 5 ? 3 : stage_request == 6 ? 2 : stage_request == 7 ? 1 : 0;
      if (threads == 256 && forced && occ[forced] > 0)
        plan.min_waves = forced;
    }
    return plan;
  }

  CoopPlan plan_coop_full(const micm_hip_mechanism* m, int stages, int stage_request, int threads_request,
                          int lu_blocking)
  {
    CoopPlan plan;
    const int threads = coop_threads_of(threads_request);
    const DevCoop c = coop_with_blocking(m, lu_blocking, threads);
    if (!m->t.coop_levels_ok)
      return plan;
    constexpr size_t kLdsPerCu = 160 * 1024;
    // the product buffer of the two-phase back substitution
    const int p_solve = c.bwd.max_level_terms;
    const size_t base = coop_lds_doubles(c.nnz_alu, m->dev.nnz_jac, m->dev.n_species, stages, m->dev.n_reactions,
                                         c.n_terms, p_solve) *
                        sizeof(double);
    const size_t solve = (coop_solve_table_words(c) + 4) * sizeof(uint32_t);
    const size_t lu = (coop_lu_table_words(c) + 2) * sizeof(uint32_t);
    const size_t limit = lds_limit();
    if (base > limit)
      return plan;
    int stage = 0;
    if (stage_request >= 0)
    {
      stage = stage_request & 3;
    }
    else
    {
      if (base + solve <= limit)
        stage |= 1;
      const size_t with_solve = base + ((stage & 1) ? solve : 0);
      if (with_solve + lu <= limit && kLdsPerCu / (with_solve + lu) >= kLdsPerCu / with_solve)
        stage |= 2;
    }
    // a requested staging that does not fit drops the LU tables first, then the rest
    if (base + ((stage & 1) ? solve : 0) + ((stage & 2) ? lu : 0) > limit)
      stage &= ~2;
    if (base + ((stage & 1) ? solve : 0) > limit)
      stage &= ~1;
    plan.lds = base + ((stage & 1) ? solve : 0) + ((stage & 2) ? lu : 0);
    plan.stage = stage;
    plan.p_capacity = p_solve;
    plan.threads = threads;
    return plan;
  }

  size_t coop_lds_bytes(const micm_hip_mechanism* m, int stages)
  {
    return plan_coop(m, stages).lds;
  }

  // Workgroups of one cooperative launch: as many as fit every CU.
  int coop_blocks(const CoopPlan& plan, int64_t& blocks)
  {
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    blocks = static_cast<int64_t>(cus) *
             std::max(coop_occupancy(plan.threads, plan.stage, plan.compact, plan.lds, plan.min_waves), 1);
    return MICM_HIP_OK;
  }

  template<int NT, int STAGE, bool COMPACT, int MINW, bool RECORD>
  int launch_coop_build(micm_hip_state* s, const DevParams& p, const CoopPlan& plan, const int32_t* list,
                        const int* count, int64_t n_fixed, const CellControl* resume, bool update_rates, hipStream_t st)
  {
    const micm_hip_mechanism* m = s->mech;
    const void* fn = reinterpret_cast<const void*>(k_coop<NT, STAGE, COMPACT, MINW, RECORD>);
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(plan.lds)));
    int64_t blocks = 0;
    if (int rc = coop_blocks(plan, blocks))
      return rc;
    // the queue counter and the compact layout's rate-constant rows: the state's own, or
    // the no-allocation solve's static scratch
    const bool ext = s->ext_next != nullptr;
    int* next = ext ? s->ext_next : s->coop_next;
    double* k_rows = ext ? s->ext_k : s->coop_k;
    if (COMPACT && (ext ? s->ext_k_rows : s->coop_k_rows) < blocks)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "cooperative scratch not prepared");
    if (!list)
      blocks = std::max<int64_t>(1, std::min(blocks, n_fixed));
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, next);
    hipLaunchKernelGGL((k_coop<NT, STAGE, COMPACT, MINW, RECORD>), dim3(static_cast<unsigned>(blocks)), dim3(NT),
                       plan.lds, st, m->dev, coop_with_blocking(m, s->lu_blocking, NT), p, plan.p_capacity, n_fixed, list,
                       count,
                       next, resume, s->y, s->k, s->atol,
                       s->status, s->stats, s->final_time, update_rates ? s->cond : nullptr,
                       update_rates ? s->custom : nullptr, k_rows, s->recording);
    HIP_TRY(hipGetLastError());
    return MICM_HIP_OK;
  }

  template<int NT, int STAGE, bool COMPACT = false, int MINW = coop_min_waves_per_eu<NT>()>
  int launch_coop_shape(micm_hip_state* s, const DevParams& p, const CoopPlan& plan, const int32_t* list,
                        const int* count, int64_t n_fixed, const CellControl* resume, bool update_rates, hipStream_t st)
  {
    if (s->recording.records)
      return launch_coop_build<NT, STAGE, COMPACT, MINW, true>(s, p, plan, list, count, n_fixed, resume, update_rates,
                                                             st);
    return launch_coop_build<NT, STAGE, COMPACT, MINW, false>(s, p, plan, list, count, n_fixed, resume, update_rates,
                                                            st);
  }

  template<int NT>
  int launch_coop_nt(micm_hip_state* s, const DevParams& p, const CoopPlan& plan, const int32_t* list,
                     const int* count, int64_t n_fixed, const CellControl* resume, bool update_rates, hipStream_t st)
  {
    if constexpr (NT == 256)
    {
      if (plan.compact && plan.min_waves == 1)
        return launch_coop_shape<NT, 0, true, 1>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
      if (plan.compact && plan.min_waves == 2)
        return launch_coop_shape<NT, 0, true, 2>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
    }
    if (plan.compact)
      return launch_coop_shape<NT, 0, true>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
    switch (plan.stage)
    {
      case 1: return launch_coop_shape<NT, 1>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
      case 2: return launch_coop_shape<NT, 2>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
      case 3: return launch_coop_shape<NT, 3>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
      default: return launch_coop_shape<NT, 0>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
    }
  }

  int launch_coop(micm_hip_state* s, const DevParams& p, const int32_t* list, const int* count, int64_t n_fixed,
                  const CellControl* resume, bool update_rates, hipStream_t st)
  {
    const CoopPlan plan = plan_coop(s->mech, p.stages, s->coop_stage, s->coop_threads, s->lu_blocking);
    if (plan.lds == 0)
      return fail(MICM_HIP_ERR_UNSUPPORTED, "the cooperative form does not fit this mechanism in LDS");
    switch (plan.threads)
    {
      case 128: return launch_coop_nt<128>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
      case 256: return launch_coop_nt<256>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
      case 512: return launch_coop_nt<512>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
      case 1024: return launch_coop_nt<1024>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
      default: return launch_coop_nt<kWave>(s, p, plan, list, count, n_fixed, resume, update_rates, st);
    }
  }

  // Device memory a per-cell form needs beyond the state's arrays, allocated before any
  // launch (and outside any stream capture).
  int persistent_blocks(const micm_hip_state* s, int64_t& blocks)
  {
    int dev = 0, cus = 0, blocks_per_cu = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, k_persistent, kBlock, 0));
    if (blocks_per_cu < 1)
      blocks_per_cu = 1;
    blocks = static_cast<int64_t>(cus) * blocks_per_cu;
    const int64_t needed = (s->n_groups + (kBlock / kGroup) - 1) / (kBlock / kGroup);
    blocks = std::max<int64_t>(1, std::min(blocks, needed));
    return MICM_HIP_OK;
  }

  int prepare_solve(micm_hip_state* s, int32_t form, int stages)
  {
    if (form == MICM_HIP_SOLVE_COOPERATIVE || form == MICM_HIP_SOLVE_PERSISTENT)
    {
      const CoopPlan plan = plan_coop(s->mech, stages, s->coop_stage, s->coop_threads, s->lu_blocking);
      if (plan.compact)
      {
        int64_t rows = 0;
        if (int rc = coop_blocks(plan, rows))
          return rc;
        if (s->coop_k_rows < rows)
        {
          if (s->coop_k)
            HIP_TRY(hipFree(s->coop_k));
          s->coop_k = nullptr;
          s->coop_k_rows = 0;
          HIP_TRY(hipMalloc(&s->coop_k, sizeof(double) * static_cast<size_t>(rows) *
                                            static_cast<size_t>(std::max(s->mech->dev.n_reactions, 1))));
          s->coop_k_rows = rows;
        }
      }
    }
    if (form != MICM_HIP_SOLVE_PERSISTENT)
      return MICM_HIP_OK;
    int64_t blocks = 0;
    if (int rc = persistent_blocks(s, blocks))
      return rc;
    return ensure_slots(s, blocks * (kBlock / kGroup), stages);
  }

  // The launches of one Solve of a per-cell form (persistent or cooperative), enqueued
  // on `st` without waiting; prepare_solve must have run.
  // The processing order of the next Solve from this Solve's per-cell attempt counts.
  // -J of the cells marked kJacobianDue, formed from zero by the LDS-tiled gather
  int launch_jacobian_due(const micm_hip_mechanism* m, int64_t n_cells, const double* k, const double* y, double* jac,
                          const CellControl* ctl, hipStream_t st)
  {
    const int tc = gather_tile_cells(m);
    const size_t lds = gather_tile_bytes(m, tc);
    switch (tc)
    {
      case 16: LAUNCH_GATHER(k_jacobian_due_tiled, 16, lds, st, m->dev, m->coop, n_cells, k, y, jac, ctl); break;
      case 8: LAUNCH_GATHER(k_jacobian_due_tiled, 8, lds, st, m->dev, m->coop, n_cells, k, y, jac, ctl); break;
      case 4: LAUNCH_GATHER(k_jacobian_due_tiled, 4, lds, st, m->dev, m->coop, n_cells, k, y, jac, ctl); break;
      default: LAUNCH_GATHER(k_jacobian_due_tiled, 2, lds, st, m->dev, m->coop, n_cells, k, y, jac, ctl); break;
    }
    HIP_TRY(hipGetLastError());
    return MICM_HIP_OK;
  }

  int enqueue_order(micm_hip_state* s, hipStream_t st)
  {
    const unsigned grid = static_cast<unsigned>((s->n_cells + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_order_clear, dim3(1), dim3(kBlock), 0, st, s->order_bins);
    hipLaunchKernelGGL(k_order_count, dim3(grid), dim3(kBlock), 0, st, s->stats, s->n_cells, s->order_bins);
    hipLaunchKernelGGL(k_order_offsets, dim3(1), dim3(1), 0, st, s->order_bins);
    hipLaunchKernelGGL(k_order_place, dim3(grid), dim3(kBlock), 0, st, s->stats, s->n_cells, s->order_bins, s->order);
    HIP_TRY(hipGetLastError());
    s->has_order = true;
    return MICM_HIP_OK;
  }

  int enqueue_solve_cells(micm_hip_state* s, const DevParams& p, int32_t form, bool update_rates, hipStream_t st);

  // A per-cell Solve (cells in the order of the previous Solve's attempt counts, most
  // first, once there was one), then the order for the next.
  int enqueue_solve(micm_hip_state* s, const DevParams& p, int32_t form, bool update_rates, hipStream_t st)
  {
    if (int rc = enqueue_solve_cells(s, p, form, update_rates, st))
      return rc;
    return enqueue_order(s, st);
  }

  int enqueue_solve_cells(micm_hip_state* s, const DevParams& p, int32_t form, bool update_rates, hipStream_t st)
  {
    const micm_hip_mechanism* m = s->mech;
    const DevMech& d = m->dev;
    if (form == MICM_HIP_SOLVE_COOPERATIVE)
      return s->has_order ? launch_coop(s, p, s->order, s->order_bins + kOrderBins, 0, nullptr, update_rates, st)
                          : launch_coop(s, p, nullptr, nullptr, s->n_cells, nullptr, update_rates, st);
    int64_t blocks = 0;
    if (int rc = persistent_blocks(s, blocks))
      return rc;
    if (s->n_slots < blocks * (kBlock / kGroup))
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "persistent scratch not prepared");
    Work w = slot_work(s);
    DevParams pp = p;
    const bool coop_ok = coop_lds_bytes(m, p.stages) > 0;
    pp.handoff_steps = !coop_ok ? 0 : (s->handoff_steps >= 0 ? s->handoff_steps : 48);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, s->queue_next);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, s->handoff_count);
    Queue q{ s->queue_next, s->has_order ? s->order : nullptr };
    hipLaunchKernelGGL(k_persistent, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, st, d, pp, s->n_cells, q, w,
                       s->y, s->k, s->atol, s->status, s->stats, s->final_time, s->ctl, s->handoff_list,
                       s->handoff_count, update_rates ? s->cond : nullptr, update_rates ? s->custom : nullptr);
    HIP_TRY(hipGetLastError());
    // handed-off cells carry the rate constants the persistent kernel stored
    if (pp.handoff_steps > 0)
      if (int rc = launch_coop(s, p, s->handoff_list, s->handoff_count, 0, s->ctl, false, st))
        return rc;
    return MICM_HIP_OK;
  }
}  // namespace

extern "C"
{
  const char* micm_hip_last_error(void)
  {
    return g_last_error.c_str();
  }

  int micm_hip_rosenbrock_parameters_preset(int32_t preset, micm_hip_rosenbrock_parameters* out)
  {
    if (!out)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null output");
    if (!rosenbrock_preset(preset, *out))
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "unknown Rosenbrock preset " + std::to_string(preset));
    return MICM_HIP_OK;
  }

  int micm_hip_mechanism_create(const micm_hip_mechanism_desc* desc, micm_hip_mechanism** out)
  {
    if (!desc || !out)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    auto m = std::make_unique<micm_hip_mechanism>();
    try
    {
      m->t = compile_mechanism(*desc);
    }
    catch (const std::domain_error& e)
    {
      return fail(MICM_HIP_ERR_SPECIES, e.what());
    }
    catch (const std::exception& e)
    {
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, e.what());
    }
    {
      const MechanismTables& t = m->t;
      DevMech& d = m->dev;
      d.n_species = static_cast<int32_t>(t.n_species);
      d.n_reactions = static_cast<int32_t>(t.n_reactions);
      d.n_custom = static_cast<int32_t>(t.n_custom);
      d.nnz_jac = static_cast<int32_t>(t.nnz_jacobian());
      d.nnz_lower = static_cast<int32_t>(t.nnz_lower());
      d.nnz_upper = static_cast<int32_t>(t.nnz_upper());
      d.n_terms = static_cast<int32_t>(t.term_process.size());
      d.n_lu_elements = static_cast<int32_t>(t.lu_out.size());
    }
    *out = m.release();
    return MICM_HIP_OK;
  }

  void micm_hip_mechanism_destroy(micm_hip_mechanism* m)
  {
    if (!m)
      return;
    if (m->blob)
      (void)hipFree(m->blob);
    delete m;
  }

  int micm_hip_mechanism_get_info(const micm_hip_mechanism* m, micm_hip_mechanism_info* out)
  {
    if (int rc = check_mech(m))
      return rc;
    if (!out)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null output");
    const auto& t = m->t;
    out->n_species = t.n_species;
    out->n_reactions = t.n_reactions;
    out->n_custom_parameters = t.n_custom;
    out->nnz_jacobian = t.nnz_jacobian();
    out->nnz_lower = t.nnz_lower();
    out->nnz_upper = t.nnz_upper();
    out->n_process_infos = static_cast<int64_t>(t.term_process.size());
    out->n_jacobian_terms = static_cast<int64_t>(t.term_flat.size());
    out->lu_multiply_adds = t.lu_multiply_adds();
    out->n_reactant_slots = static_cast<int64_t>(t.reactant_ids.size());
    out->n_product_slots = static_cast<int64_t>(t.product_ids.size());
    return MICM_HIP_OK;
  }

  int micm_hip_mechanism_get_species_order(const micm_hip_mechanism* m, int64_t* out)
  {
    if (int rc = check_mech(m))
      return rc;
    if (!out)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null output");
    std::copy(m->t.state_to_declared.begin(), m->t.state_to_declared.end(), out);
    return MICM_HIP_OK;
  }

  int micm_hip_mechanism_get_reaction_order(const micm_hip_mechanism* m, int64_t* out)
  {
    if (int rc = check_mech(m))
      return rc;
    if (!out)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null output");
    std::copy(m->t.sorted_to_declared.begin(), m->t.sorted_to_declared.end(), out);
    return MICM_HIP_OK;
  }

  int micm_hip_mechanism_get_custom_parameter_owner(const micm_hip_mechanism* m, int64_t* out)
  {
    if (int rc = check_mech(m))
      return rc;
    if (!out && !m->t.custom_owner.empty())
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null output");
    std::copy(m->t.custom_owner.begin(), m->t.custom_owner.end(), out);
    return MICM_HIP_OK;
  }

  int micm_hip_mechanism_get_patterns(const micm_hip_mechanism* m, int64_t* jr, int64_t* jc, int64_t* lr, int64_t* lc,
                                      int64_t* ur, int64_t* uc)
  {
    if (int rc = check_mech(m))
      return rc;
    auto dump = [](const std::vector<int32_t>& ptr, const std::vector<int32_t>& col, int64_t* rows, int64_t* cols)
    {
      for (size_t i = 0; i + 1 < ptr.size(); ++i)
        for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p)
        {
          if (rows)
            rows[p] = static_cast<int64_t>(i);
          if (cols)
            cols[p] = col[p];
        }
    };
    dump(m->t.jac_row_ptr, m->t.jac_col, jr, jc);
    dump(m->t.lower_row_ptr, m->t.lower_col, lr, lc);
    dump(m->t.upper_row_ptr, m->t.upper_col, ur, uc);
    return MICM_HIP_OK;
  }

  int micm_hip_mechanism_get_jacobian_diagonal(const micm_hip_mechanism* m, int64_t* out)
  {
    if (int rc = check_mech(m))
      return rc;
    if (!out)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null output");
    std::copy(m->t.jac_diag.begin(), m->t.jac_diag.end(), out);
    return MICM_HIP_OK;
  }

  // ---- kernel entry points

#define KERNEL_PRELUDE()                                                   \
  if (int rc = check_device_mech(m))                                       \
    return rc;                                                             \
  if (n_cells < 0)                                                         \
    return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "negative cell count");     \
  if (n_cells == 0)                                                        \
    return MICM_HIP_OK;                                                    \
  hipStream_t st = static_cast<hipStream_t>(stream)

  int micm_hip_calculate_rate_constants(const micm_hip_mechanism* m, int64_t n_cells, const double* cond,
                                        const double* custom, double* k, void* stream)
  {
    KERNEL_PRELUDE();
    if (!cond || !k || (m->dev.n_custom > 0 && !custom))
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null device pointer");
    hipLaunchKernelGGL(k_rate_constants, rate_grid(n_cells, m->dev.n_reactions), dim3(kBlock), 0, st, m->dev, n_cells,
                       cond, custom, k);
    HIP_TRY(hipGetLastError());
    return MICM_HIP_OK;
  }

  // (C++ linkage: templates; the entry points below are the C-ABI)
  extern "C++" template<bool FROM_ZERO>
  int add_forcing_entry(const micm_hip_mechanism* m, int64_t n_cells, const double* k, const double* y, double* f,
                        void* stream)
  {
    KERNEL_PRELUDE();
    if (!k || !y || !f)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null device pointer");
    const int tc = gather_tile_cells(m);
    if (tc > 0)
    {
      const size_t lds = gather_tile_bytes(m, tc);
      switch (tc)
      {
        case 16: LAUNCH_GATHER_Z(k_add_forcing_tiled, 16, FROM_ZERO, lds, st, m->dev, m->coop, n_cells, k, y, f); break;
        case 8: LAUNCH_GATHER_Z(k_add_forcing_tiled, 8, FROM_ZERO, lds, st, m->dev, m->coop, n_cells, k, y, f); break;
        case 4: LAUNCH_GATHER_Z(k_add_forcing_tiled, 4, FROM_ZERO, lds, st, m->dev, m->coop, n_cells, k, y, f); break;
        default: LAUNCH_GATHER_Z(k_add_forcing_tiled, 2, FROM_ZERO, lds, st, m->dev, m->coop, n_cells, k, y, f); break;
      }
    }
    else
    {
      hipLaunchKernelGGL(k_add_forcing<FROM_ZERO>, dim3(grid_for(n_cells)), dim3(kBlock), 0, st, m->dev, n_cells, k, y,
                         f);
    }
    HIP_TRY(hipGetLastError());
    return MICM_HIP_OK;
  }

  extern "C++" template<bool FROM_ZERO>
  int subtract_jacobian_entry(const micm_hip_mechanism* m, int64_t n_cells, const double* k, const double* y,
                              double* jac, void* stream)
  {
    KERNEL_PRELUDE();
    if (!k || !y || !jac)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null device pointer");
    const int tc = gather_tile_cells(m);
    if (tc > 0)
    {
      const size_t lds = gather_tile_bytes(m, tc);
      switch (tc)
      {
        case 16:
          LAUNCH_GATHER_Z(k_subtract_jacobian_tiled, 16, FROM_ZERO, lds, st, m->dev, m->coop, n_cells, k, y, jac);
          break;
        case 8:
          LAUNCH_GATHER_Z(k_subtract_jacobian_tiled, 8, FROM_ZERO, lds, st, m->dev, m->coop, n_cells, k, y, jac);
          break;
        case 4:
          LAUNCH_GATHER_Z(k_subtract_jacobian_tiled, 4, FROM_ZERO, lds, st, m->dev, m->coop, n_cells, k, y, jac);
          break;
        default:
          LAUNCH_GATHER_Z(k_subtract_jacobian_tiled, 2, FROM_ZERO, lds, st, m->dev, m->coop, n_cells, k, y, jac);
          break;
      }
    }
    else
    {
      hipLaunchKernelGGL(k_subtract_jacobian<FROM_ZERO>, dim3(grid_for(n_cells)), dim3(kBlock), 0, st, m->dev, n_cells, k,
                         y, jac);
    }
    HIP_TRY(hipGetLastError());
    return MICM_HIP_OK;
  }

  int micm_hip_add_forcing_terms(const micm_hip_mechanism* m, int64_t n_cells, const double* k, const double* y,
                                 double* f, void* stream)
  {
    return add_forcing_entry<false>(m, n_cells, k, y, f, stream);
  }

  int micm_hip_fill_and_add_forcing_terms(const micm_hip_mechanism* m, int64_t n_cells, const double* k,
                                          const double* y, double* f, void* stream)
  {
    return add_forcing_entry<true>(m, n_cells, k, y, f, stream);
  }

  int micm_hip_subtract_jacobian_terms(const micm_hip_mechanism* m, int64_t n_cells, const double* k, const double* y,
                                       double* jac, void* stream)
  {
    return subtract_jacobian_entry<false>(m, n_cells, k, y, jac, stream);
  }

  int micm_hip_fill_and_subtract_jacobian_terms(const micm_hip_mechanism* m, int64_t n_cells, const double* k,
                                                const double* y, double* jac, void* stream)
  {
    return subtract_jacobian_entry<true>(m, n_cells, k, y, jac, stream);
  }

  int micm_hip_alpha_minus_jacobian(const micm_hip_mechanism* m, int64_t n_cells, double* jac, double alpha, void* stream)
  {
    KERNEL_PRELUDE();
    if (!jac)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null device pointer");
    hipLaunchKernelGGL(k_alpha_minus_jacobian, dim3(grid_for(n_cells)), dim3(kBlock), 0, st, m->dev, n_cells, jac, alpha);
    HIP_TRY(hipGetLastError());
    return MICM_HIP_OK;
  }

  int micm_hip_lu_decompose(const micm_hip_mechanism* m, int64_t n_cells, const double* jac, double* lower, double* upper,
                            void* stream)
  {
    KERNEL_PRELUDE();
    if (!jac || !lower || !upper)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null device pointer");
    const int tc = tile_cells(m, static_cast<size_t>(m->coop.nnz_alu));
    if (tc > 0)
    {
      const size_t lds = static_cast<size_t>(m->coop.nnz_alu) * tc * sizeof(double);
      const unsigned blocks = tiled_blocks(tc, n_cells);
      switch (tc)
      {
        case 16: LAUNCH_TILED(k_lu_decompose_tiled<16>, blocks, lds, st, m->dev, m->coop, n_cells, jac, lower, upper); break;
        case 8: LAUNCH_TILED(k_lu_decompose_tiled<8>, blocks, lds, st, m->dev, m->coop, n_cells, jac, lower, upper); break;
        case 4: LAUNCH_TILED(k_lu_decompose_tiled<4>, blocks, lds, st, m->dev, m->coop, n_cells, jac, lower, upper); break;
        case 2: LAUNCH_TILED(k_lu_decompose_tiled<2>, blocks, lds, st, m->dev, m->coop, n_cells, jac, lower, upper); break;
        default: LAUNCH_TILED(k_lu_decompose_tiled<1>, blocks, lds, st, m->dev, m->coop, n_cells, jac, lower, upper); break;
      }
    }
    else
    {
      hipLaunchKernelGGL(k_lu_decompose, dim3(grid_for(n_cells)), dim3(kBlock), 0, st, m->dev, n_cells, jac, lower, upper);
    }
    HIP_TRY(hipGetLastError());
    return MICM_HIP_OK;
  }

  int micm_hip_lu_solve(const micm_hip_mechanism* m, int64_t n_cells, const double* lower, const double* upper, double* x,
                        void* stream)
  {
    KERNEL_PRELUDE();
    if (!lower || !upper || !x)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null device pointer");
    const int tc = tile_cells(m, static_cast<size_t>(m->coop.nnz_alu) + 2 * static_cast<size_t>(m->dev.n_species));
    if (tc > 0)
    {
      size_t lds = (static_cast<size_t>(m->coop.nnz_alu) + 2 * static_cast<size_t>(m->dev.n_species)) * tc *
                   sizeof(double);
      // one cell per workgroup: a product buffer for the two-phase back substitution
      // from what two workgroups per CU leave (levels with more terms run task by task)
      int p_capacity = 0;
      if (tc == 1)
      {
        const size_t spare = lds_limit() / 2 > lds ? (lds_limit() / 2 - lds) / sizeof(double) : 0;
        p_capacity = static_cast<int>(std::min<size_t>(static_cast<size_t>(m->coop.bwd.max_level_terms), spare));
        lds += static_cast<size_t>(p_capacity) * sizeof(double);
      }
      const unsigned blocks = tiled_blocks(tc, n_cells);
      switch (tc)
      {
        case 16:
          LAUNCH_TILED(k_lu_solve_tiled<16>, blocks, lds, st, m->dev, m->coop, n_cells, lower, upper, x, p_capacity);
          break;
        case 8:
          LAUNCH_TILED(k_lu_solve_tiled<8>, blocks, lds, st, m->dev, m->coop, n_cells, lower, upper, x, p_capacity);
          break;
        case 4:
          LAUNCH_TILED(k_lu_solve_tiled<4>, blocks, lds, st, m->dev, m->coop, n_cells, lower, upper, x, p_capacity);
          break;
        case 2:
          LAUNCH_TILED(k_lu_solve_tiled<2>, blocks, lds, st, m->dev, m->coop, n_cells, lower, upper, x, p_capacity);
          break;
        default:
          LAUNCH_TILED(k_lu_solve_tiled<1>, blocks, lds, st, m->dev, m->coop, n_cells, lower, upper, x, p_capacity);
          break;
      }
    }
    else
    {
      hipLaunchKernelGGL(k_lu_solve, dim3(grid_for(n_cells)), dim3(kBlock), 0, st, m->dev, n_cells, lower, upper, x);
    }
    HIP_TRY(hipGetLastError());
    return MICM_HIP_OK;
  }

  int micm_hip_normalized_error(const micm_hip_mechanism* m, int64_t n_cells, const double* y, const double* ynew,
                                const double* yerr, const double* atol, double rtol, double* error, void* stream)
  {
    KERNEL_PRELUDE();
    if (!y || !ynew || !yerr || !atol || !error)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null device pointer");
    hipLaunchKernelGGL(k_normalized_error, dim3(grid_for(n_cells)), dim3(kBlock), 0, st, m->dev.n_species, n_cells, y,
                       ynew, yerr, atol, rtol, error);
    HIP_TRY(hipGetLastError());
    return MICM_HIP_OK;
  }
  int micm_hip_normalized_error_state(const micm_hip_mechanism* m, int64_t n_cells, const double* y,
                                      const double* ynew, const double* errors, const double* atol, double rtol,
                                      int64_t vector_length, double* error, void* stream)
  {
    if (int rc = check_device_mech(m))
      return rc;
    if (n_cells < 0 || vector_length < 1)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "negative cell count or vector length below 1");
    if (!error || (n_cells > 0 && (!y || !ynew || !errors || !atol)))
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null buffer");
    hipLaunchKernelGGL(k_state_error, dim3(1), dim3(kLockstepErrorThreads), 0, static_cast<hipStream_t>(stream),
                       m->dev.n_species, n_cells, vector_length, y, ynew, errors, atol, rtol, error);
    HIP_TRY(hipGetLastError());
    return MICM_HIP_OK;
  }

  int micm_hip_state_set_lockstep_order(micm_hip_state* s, int64_t vector_length)
  {
    if (!s || vector_length < 1)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null state or vector length below 1");
    s->lockstep_group = vector_length;
    return MICM_HIP_OK;
  }


  // ---- state

  int micm_hip_state_create(const micm_hip_mechanism* m, int64_t n_cells, micm_hip_state** out)
  {
    if (int rc = check_device_mech(m))
      return rc;
    if (!out || n_cells <= 0)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "state needs at least one cell");
    *out = nullptr;
    auto s = std::make_unique<micm_hip_state>();
    s->mech = m;
    s->n_cells = n_cells;
    s->n_groups = n_groups_of(n_cells);
    const auto& d = m->dev;
    const size_t G = static_cast<size_t>(s->n_groups) * kGroup;
    auto alloc = [&](double** p, size_t n) -> int {
      HIP_TRY(hipMalloc(p, std::max<size_t>(n, 1) * sizeof(double)));
      HIP_TRY(hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(double)));
      return MICM_HIP_OK;
    };
    int rc = 0;
    if ((rc = alloc(&s->y, G * d.n_species)) || (rc = alloc(&s->cond, G * 3)) || (rc = alloc(&s->custom, G * d.n_custom)) ||
        (rc = alloc(&s->k, G * d.n_reactions)) || (rc = alloc(&s->jac, G * d.nnz_jac)) ||
        (rc = alloc(&s->lower, G * d.nnz_lower)) || (rc = alloc(&s->upper, G * d.nnz_upper)) ||
        (rc = alloc(&s->atol, d.n_species)) || (rc = alloc(&s->f0, G * d.n_species)) ||
        (rc = alloc(&s->ynew, G * d.n_species)) || (rc = alloc(&s->yerr, G * d.n_species)) ||
        (rc = alloc(&s->K, G * d.n_species * 6)) || (rc = alloc(&s->final_time, n_cells)) ||
        (rc = alloc(&s->lockstep_error, 1)))
    {
      micm_hip_state_destroy(s.release());
      return rc;
    }
    HIP_TRY(hipMalloc(&s->ctl, sizeof(CellControl) * n_cells));
    HIP_TRY(hipMalloc(&s->status, sizeof(int32_t) * n_cells));
    HIP_TRY(hipMemset(s->status, 0, sizeof(int32_t) * n_cells));
    HIP_TRY(hipMalloc(&s->stats, sizeof(int64_t) * 7 * n_cells));
    HIP_TRY(hipMemset(s->stats, 0, sizeof(int64_t) * 7 * n_cells));
    HIP_TRY(hipMalloc(&s->active, sizeof(unsigned int)));
    HIP_TRY(hipMalloc(&s->queue_next, sizeof(int)));
    HIP_TRY(hipMalloc(&s->order, sizeof(int32_t) * n_cells));
    HIP_TRY(hipMalloc(&s->handoff_list, sizeof(int32_t) * n_cells));
    HIP_TRY(hipMalloc(&s->handoff_count, sizeof(int)));
    HIP_TRY(hipMalloc(&s->coop_next, sizeof(int)));
    HIP_TRY(hipMalloc(&s->order_bins, sizeof(int) * (kOrderBins + 1)));
    HIP_TRY(hipHostMalloc(&s->host_active, sizeof(unsigned int)));
    HIP_TRY(hipMemcpy(s->atol, m->t.absolute_tolerances.data(), sizeof(double) * d.n_species, hipMemcpyHostToDevice));
    *out = s.release();
    return MICM_HIP_OK;
  }

  void micm_hip_state_destroy(micm_hip_state* s)
  {
    if (!s)
      return;
    for (double* p : { s->y, s->cond, s->custom, s->k, s->jac, s->lower, s->upper, s->atol, s->f0, s->ynew, s->yerr, s->K,
                       s->final_time, s->lockstep_error, s->slot_blob, s->coop_k })
      if (p)
        (void)hipFree(p);
    if (s->ctl)
      (void)hipFree(s->ctl);
    if (s->status)
      (void)hipFree(s->status);
    if (s->stats)
      (void)hipFree(s->stats);
    if (s->active)
      (void)hipFree(s->active);
    if (s->queue_next)
      (void)hipFree(s->queue_next);
    if (s->order)
      (void)hipFree(s->order);
    if (s->handoff_list)
      (void)hipFree(s->handoff_list);
    if (s->handoff_count)
      (void)hipFree(s->handoff_count);
    if (s->coop_next)
      (void)hipFree(s->coop_next);
    if (s->order_bins)
      (void)hipFree(s->order_bins);
    if (s->host_active)
      (void)hipHostFree(s->host_active);
    delete s;
  }

  int micm_hip_state_get_pointers(const micm_hip_state* s, micm_hip_state_pointers* out)
  {
    if (!s || !out)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null argument");
    out->variables = s->y;
    out->conditions = s->cond;
    out->custom_parameters = s->custom;
    out->rate_constants = s->k;
    out->jacobian = s->jac;
    out->lower = s->lower;
    out->upper = s->upper;
    out->absolute_tolerances = s->atol;
    out->cell_state = s->status;
    out->cell_stats = s->stats;
    out->cell_final_time = s->final_time;
    out->n_cells = s->n_cells;
    out->n_groups = s->n_groups;
    return MICM_HIP_OK;
  }

  int micm_hip_state_set_variables(micm_hip_state* s, const double* host)
  {
    return set_grouped(s, s ? s->y : nullptr, host, s ? s->mech->dev.n_species : 0);
  }
  int micm_hip_state_get_variables(const micm_hip_state* s, double* host)
  {
    return get_grouped(s, s ? s->y : nullptr, host, s ? s->mech->dev.n_species : 0);
  }
  int micm_hip_state_set_conditions(micm_hip_state* s, const double* host)
  {
    return set_grouped(s, s ? s->cond : nullptr, host, 3);
  }
  int micm_hip_state_set_custom_parameters(micm_hip_state* s, const double* host)
  {
    return set_grouped(s, s ? s->custom : nullptr, host, s ? s->mech->dev.n_custom : 0);
  }
  int micm_hip_state_get_rate_constants(const micm_hip_state* s, double* host)
  {
    return get_grouped(s, s ? s->k : nullptr, host, s ? s->mech->dev.n_reactions : 0);
  }
  int micm_hip_state_set_rate_constants(micm_hip_state* s, const double* host)
  {
    return set_grouped(s, s ? s->k : nullptr, host, s ? s->mech->dev.n_reactions : 0);
  }

  int micm_hip_state_set_tolerances(micm_hip_state* s, double rtol, const double* atol)
  {
    if (!s)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null state");
    if (!(rtol > 0))
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "relative tolerance must be positive");
    s->rtol = rtol;
    if (atol)
    {
      HIP_TRY(hipMemcpyAsync(s->atol, atol, sizeof(double) * s->mech->dev.n_species, hipMemcpyHostToDevice, s->stream));
      HIP_TRY(hipStreamSynchronize(s->stream));
    }
    return MICM_HIP_OK;
  }

  int micm_hip_state_get_results(const micm_hip_state* s, int32_t* states, int64_t* stats, double* final_time)
  {
    if (!s)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null state");
    if (states)
      HIP_TRY(hipMemcpyAsync(states, s->status, sizeof(int32_t) * s->n_cells, hipMemcpyDeviceToHost, s->stream));
    if (stats)
      HIP_TRY(hipMemcpyAsync(stats, s->stats, sizeof(int64_t) * 7 * s->n_cells, hipMemcpyDeviceToHost, s->stream));
    if (final_time)
      HIP_TRY(hipMemcpyAsync(final_time, s->final_time, sizeof(double) * s->n_cells, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MICM_HIP_OK;
  }

  int micm_hip_state_set_stream(micm_hip_state* s, void* stream)
  {
    if (!s)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null state");
    s->stream = static_cast<hipStream_t>(stream);
    return MICM_HIP_OK;
  }

  int micm_hip_state_set_handoff(micm_hip_state* s, int64_t steps)
  {
    if (!s)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null state");
    s->handoff_steps = steps;
    return MICM_HIP_OK;
  }

  int micm_hip_state_set_cooperative_shape(micm_hip_state* s, int32_t threads, int32_t stage)
  {
    if (!s)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null state");
    if (threads != 0 && threads != 64 && threads != 128 && threads != 256 && threads != 512 && threads != 1024)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "threads per cell must be 0, 64, 128, 256, 512 or 1024");
    if (stage < -1 || stage > 7)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "stage must be -1 .. 7");
    s->coop_threads = threads;
    s->coop_stage = stage;
    return MICM_HIP_OK;
  }

  int micm_hip_state_set_lu_blocking(micm_hip_state* s, int32_t min_segment)
  {
    if (!s)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null state");
    if (min_segment != 0 && lu_blocking_index(min_segment) < 0)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "LU blocking must be 0, -1, 1, 2, 4 or 8");
    s->lu_blocking = min_segment;
    return MICM_HIP_OK;
  }

  int micm_hip_mechanism_cooperative_supported(const micm_hip_mechanism* m, int32_t stages, int32_t* out)
  {
    if (int rc = check_mech(m))
      return rc;
    if (!out || stages < 1 || stages > 6)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "bad argument");
    *out = coop_lds_bytes(m, stages) > 0 ? 1 : 0;
    return MICM_HIP_OK;
  }

  int micm_hip_update_state_parameters(const micm_hip_mechanism* m, micm_hip_state* s)
  {
    if (int rc = check_mech(m))
      return rc;
    if (!s || s->mech != m)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "state does not belong to this mechanism");
    hipLaunchKernelGGL(k_rate_constants, rate_grid(s->n_cells, m->dev.n_reactions), dim3(kBlock), 0, s->stream, m->dev,
                       s->n_cells, s->cond, s->custom, s->k);
    HIP_TRY(hipGetLastError());
    return MICM_HIP_OK;
  }

  int micm_hip_solve(const micm_hip_mechanism* m, micm_hip_state* s, double time_step,
                     const micm_hip_rosenbrock_parameters* params, int32_t form)
  {
    if (int rc = check_mech(m))
      return rc;
    if (!s || s->mech != m || !params)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "bad state or parameters");
    if (params->stages < 1 || params->stages > 6)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "stages must be 1..6");
    DevParams p = make_params(*params, time_step, s->rtol);
    const DevMech& d = m->dev;
    hipStream_t st = s->stream;
    const unsigned grid = grid_for(s->n_cells);
    const bool update_rates = (form & MICM_HIP_SOLVE_UPDATE_RATES) != 0;
    form &= ~MICM_HIP_SOLVE_UPDATE_RATES;
    if (form == MICM_HIP_SOLVE_PLAIN || form == MICM_HIP_SOLVE_LOCKSTEP)
    {
      if (update_rates)
      {
        hipLaunchKernelGGL(k_rate_constants, rate_grid(s->n_cells, d.n_reactions), dim3(kBlock), 0, st, d, s->n_cells,
                           s->cond, s->custom, s->k);
        HIP_TRY(hipGetLastError());
      }
      const bool lockstep = form == MICM_HIP_SOLVE_LOCKSTEP;
      p.error_group = s->lockstep_group;
      Work w = state_work(s);
      hipLaunchKernelGGL(k_plain_init, dim3(grid), dim3(kBlock), 0, st, p, s->n_cells, s->ctl);
      HIP_TRY(hipGetLastError());
      // every attempt either finishes a cell or advances its step counter, and a cell
      // stops after max_number_of_steps + 1 attempts, so this bound is never reached
      const int64_t max_iterations = (params->max_number_of_steps + 2) * 64;
      // the factorisation and substitutions as the LDS-tiled level-schedule kernels when
      // the mechanism has them, else within the per-lane kernels
      const bool tiled = tile_cells(m, static_cast<size_t>(m->coop.nnz_alu)) > 0 &&
                         tile_cells(m, static_cast<size_t>(m->coop.nnz_alu) + 2 * static_cast<size_t>(d.n_species)) > 0;
      // the Jacobian as the tiled gather for the cells that start a step
      const bool gather = gather_tile_cells(m) > 0;
      // kernels with one thread per (cell, species)
      const unsigned pair_grid = static_cast<unsigned>(
          (n_groups_of(s->n_cells) * static_cast<int64_t>(d.n_species) * kGroup + kBlock - 1) / kBlock);
      // a small State of a small mechanism: the whole lockstep loop in one workgroup
      const bool fused = lockstep && s->n_cells <= kFusedLockstepMaxCells &&
                         static_cast<int64_t>(d.nnz_lower) + d.nnz_upper - d.n_species <= kFusedLockstepMaxFactor;
      if (fused)
      {
        hipLaunchKernelGGL(k_lockstep_fused, dim3(1), dim3(kFusedLockstepThreads), 0, st, d, p, s->n_cells, w, s->ctl,
                           s->atol, max_iterations);
        HIP_TRY(hipGetLastError());
      }
      for (int64_t it = 0; !fused && it < max_iterations; ++it)
      {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, st, s->active);
        if (tiled)
        {
          if (gather)
          {
            hipLaunchKernelGGL(k_plain_begin, dim3(grid), dim3(kBlock), 0, st, d, p, s->n_cells, w, s->ctl);
            if (int rc = launch_jacobian_due(m, s->n_cells, w.k, w.y, w.jac, s->ctl, st))
              return rc;
            hipLaunchKernelGGL(k_plain_alpha, dim3(grid), dim3(kBlock), 0, st, d, p, s->n_cells, w, s->ctl);
          }
          else
          {
            hipLaunchKernelGGL(k_plain_prepare, dim3(grid), dim3(kBlock), 0, st, d, p, s->n_cells, w, s->ctl);
          }
          // the initial forcing formed from zero: Fill(0) and AddForcingTerms (rosenbrock.inl:98-99)
          if (int rc = micm_hip_fill_and_add_forcing_terms(m, s->n_cells, w.k, w.y, w.f0, st))
            return rc;
          if (int rc = micm_hip_lu_decompose(m, s->n_cells, w.jac, w.lower, w.upper, st))
            return rc;
          for (int stage = 0; stage < p.stages; ++stage)
          {
            double* Ks = w.K + static_cast<size_t>(stage) * w.k_stride;
            if (stage > 0 && p.new_f[stage])
            {
              hipLaunchKernelGGL(k_stage_rhs_elem<1>, dim3(pair_grid), dim3(kBlock), 0, st, d, p, s->n_cells, w,
                                 s->ctl, stage);
              // K[stage] = f(Y_stage) from zero (rosenbrock.inl:156-157)
              if (int rc = micm_hip_fill_and_add_forcing_terms(m, s->n_cells, w.k, w.ynew, Ks, st))
                return rc;
            }
            hipLaunchKernelGGL(k_stage_rhs_elem<2>, dim3(pair_grid), dim3(kBlock), 0, st, d, p, s->n_cells, w, s->ctl,
                               stage);
            if (int rc = micm_hip_lu_solve(m, s->n_cells, w.lower, w.upper, Ks, st))
              return rc;
          }
        }
        else
        {
          hipLaunchKernelGGL(k_plain_factor, dim3(grid), dim3(kBlock), 0, st, d, p, s->n_cells, w, s->ctl);
          for (int stage = 0; stage < p.stages; ++stage)
            hipLaunchKernelGGL(k_plain_stage, dim3(grid), dim3(kBlock), 0, st, d, p, s->n_cells, w, s->ctl, stage);
        }
        hipLaunchKernelGGL(k_combine_elem, dim3(pair_grid), dim3(kBlock), 0, st, d, p, s->n_cells, w, s->ctl);
        if (!lockstep)
          hipLaunchKernelGGL(k_plain_finish, dim3(grid), dim3(kBlock), 0, st, d, p, s->n_cells, w, s->ctl, s->atol,
                             s->active);
        else
        {
          hipLaunchKernelGGL(k_lockstep_error, dim3(1), dim3(kLockstepErrorThreads), 0, st, d, p, s->n_cells, w,
                             s->ctl, s->atol, s->lockstep_error);
          hipLaunchKernelGGL(k_lockstep_control, dim3(grid), dim3(kBlock), 0, st, d, p, s->n_cells, w, s->ctl,
                             s->lockstep_error, s->active);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(s->host_active, s->active, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (*s->host_active == 0)
          break;
      }
      hipLaunchKernelGGL(k_finalize, dim3(grid), dim3(kBlock), 0, st, d.n_species, s->n_cells, s->y, s->ctl, s->status,
                         s->stats, s->final_time);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(st));
      return MICM_HIP_OK;
    }
    if (form == MICM_HIP_SOLVE_PERSISTENT || form == MICM_HIP_SOLVE_COOPERATIVE)
    {
      if (int rc = prepare_solve(s, form, p.stages))
        return rc;
      if (int rc = enqueue_solve(s, p, form, update_rates, st))
        return rc;
      HIP_TRY(hipStreamSynchronize(st));
      return MICM_HIP_OK;
    }
    return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "unknown solve form");
  }

  int micm_hip_solve_noalloc(const micm_hip_mechanism* m, micm_hip_state* s, double time_step,
                             const micm_hip_rosenbrock_parameters* params, int32_t flags)
  {
    if (int rc = check_mech(m))
      return rc;
    if (!s || s->mech != m || !params)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "bad state or parameters");
    if ((flags & ~MICM_HIP_SOLVE_UPDATE_RATES) != 0)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "flags: only MICM_HIP_SOLVE_UPDATE_RATES");
    if (params->stages < 1 || params->stages > 6)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "stages must be 1..6");
    const CoopPlan plan = plan_coop(m, static_cast<int>(params->stages), s->coop_stage, s->coop_threads, s->lu_blocking);
    if (plan.lds == 0)
      return fail(MICM_HIP_ERR_UNSUPPORTED, "the cooperative form does not fit this mechanism in LDS");
    int64_t rows = 0;
    if (plan.compact)
      if (int rc = coop_blocks(plan, rows))
        return rc;
    const size_t need = kNoAllocCounterBytes + static_cast<size_t>(rows) *
                                                   static_cast<size_t>(std::max(m->dev.n_reactions, 1)) * sizeof(double);
    if (need > kNoAllocScratchBytes)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT,
                  "this launch shape needs " + std::to_string(need) + " bytes of scratch; the no-allocation solve has " +
                      std::to_string(kNoAllocScratchBytes) + " (choose a shape without the compact layout)");
    void* base = nullptr;
    if (int rc = noalloc_scratch(&base))
      return rc;
    const DevParams p = make_params(*params, time_step, s->rtol);
    s->ext_next = static_cast<int*>(base);
    s->ext_k = reinterpret_cast<double*>(static_cast<char*>(base) + kNoAllocCounterBytes);
    s->ext_k_rows = rows;
    const int rc = enqueue_solve(s, p, MICM_HIP_SOLVE_COOPERATIVE, (flags & MICM_HIP_SOLVE_UPDATE_RATES) != 0,
                                 s->stream);
    s->ext_next = nullptr;
    s->ext_k = nullptr;
    s->ext_k_rows = 0;
    return rc;
  }

  // ---------------------------------------------------------------- discrete adjoint

  int micm_hip_tape_create(const micm_hip_mechanism* m, int64_t n_cells, int64_t max_records, micm_hip_tape** out)
  {
    if (int rc = check_device_mech(m))
      return rc;
    if (!out || n_cells < 1 || max_records < 1)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "a tape needs at least one cell and one record per cell");
    auto tp = std::make_unique<micm_hip_tape>();
    tp->mech = m;
    tp->n_cells = n_cells;
    tp->max_records = max_records;
    const size_t doubles = static_cast<size_t>(n_cells) * static_cast<size_t>(max_records) *
                           (static_cast<size_t>(m->dev.n_species) + 1);
    if (hipMalloc(&tp->records, doubles * sizeof(double)) != hipSuccess)
    {
      (void)hipGetLastError();
      return fail(MICM_HIP_ERR_OUT_OF_MEMORY, "tape of " + std::to_string(doubles * sizeof(double)) + " bytes");
    }
    if (hipMalloc(&tp->counts, sizeof(int64_t) * static_cast<size_t>(n_cells)) != hipSuccess)
    {
      (void)hipGetLastError();
      (void)hipFree(tp->records);
      return fail(MICM_HIP_ERR_OUT_OF_MEMORY, "tape counters");
    }
    HIP_TRY(hipMemset(tp->counts, 0, sizeof(int64_t) * static_cast<size_t>(n_cells)));
    *out = tp.release();
    return MICM_HIP_OK;
  }

  void micm_hip_tape_destroy(micm_hip_tape* tp)
  {
    if (!tp)
      return;
    (void)hipDeviceSynchronize();
    (void)hipFree(tp->records);
    (void)hipFree(tp->counts);
    delete tp;
  }

  int micm_hip_tape_reset(micm_hip_tape* tp)
  {
    if (!tp)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null tape");
    HIP_TRY(hipMemset(tp->counts, 0, sizeof(int64_t) * static_cast<size_t>(tp->n_cells)));
    return MICM_HIP_OK;
  }

  int micm_hip_tape_get(const micm_hip_tape* tp, int64_t first_cell, int64_t n_cells, double* records,
                        int64_t* counts)
  {
    if (!tp || first_cell < 0 || n_cells < 0 || first_cell + n_cells > tp->n_cells)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null tape or cell range outside it");
    HIP_TRY(hipDeviceSynchronize());
    const size_t per_cell = static_cast<size_t>(tp->max_records) * (static_cast<size_t>(tp->mech->dev.n_species) + 1);
    if (records && n_cells)
      HIP_TRY(hipMemcpy(records, tp->records + static_cast<size_t>(first_cell) * per_cell,
                        sizeof(double) * static_cast<size_t>(n_cells) * per_cell, hipMemcpyDeviceToHost));
    if (counts && n_cells)
      HIP_TRY(hipMemcpy(counts, tp->counts + first_cell, sizeof(int64_t) * static_cast<size_t>(n_cells),
                        hipMemcpyDeviceToHost));
    return MICM_HIP_OK;
  }

  int micm_hip_tape_set(micm_hip_tape* tp, const double* records, const int64_t* counts)
  {
    if (!tp || !records || !counts)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(tp->records, records,
                      sizeof(double) * static_cast<size_t>(tp->n_cells) * static_cast<size_t>(tp->max_records) *
                          (static_cast<size_t>(tp->mech->dev.n_species) + 1),
                      hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(tp->counts, counts, sizeof(int64_t) * static_cast<size_t>(tp->n_cells), hipMemcpyHostToDevice));
    return MICM_HIP_OK;
  }

  int micm_hip_tape_overflowed(const micm_hip_tape* tp, int64_t* n_cells_overflowed)
  {
    if (!tp || !n_cells_overflowed)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<int64_t> counts(static_cast<size_t>(tp->n_cells));
    if (int rc = micm_hip_tape_get(tp, 0, tp->n_cells, nullptr, counts.data()))
      return rc;
    int64_t n = 0;
    for (int64_t c : counts)
      n += c > tp->max_records;
    *n_cells_overflowed = n;
    return MICM_HIP_OK;
  }

  int micm_hip_solve_recorded(const micm_hip_mechanism* m, micm_hip_state* s, double time_step,
                              const micm_hip_rosenbrock_parameters* params, int32_t flags, micm_hip_tape* tape)
  {
    if (int rc = check_mech(m))
      return rc;
    if (!s || s->mech != m || !params || !tape || tape->mech != m || tape->n_cells != s->n_cells)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "bad state, parameters or tape");
    if ((flags & ~MICM_HIP_SOLVE_UPDATE_RATES) != 0)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "flags: only MICM_HIP_SOLVE_UPDATE_RATES");
    if (params->stages < 1 || params->stages > 6)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "stages must be 1..6");
    if (plan_coop(m, static_cast<int>(params->stages), s->coop_stage, s->coop_threads, s->lu_blocking).lds == 0)
      return fail(MICM_HIP_ERR_UNSUPPORTED, "recorded solves run the cooperative form, which does not fit this mechanism");
    const DevParams p = make_params(*params, time_step, s->rtol);
    if (int rc = prepare_solve(s, MICM_HIP_SOLVE_COOPERATIVE, p.stages))
      return rc;
    s->recording = Tape{ tape->records, tape->counts, tape->max_records };
    const int rc = enqueue_solve(s, p, MICM_HIP_SOLVE_COOPERATIVE, (flags & MICM_HIP_SOLVE_UPDATE_RATES) != 0, s->stream);
    s->recording = Tape{ nullptr, nullptr, 0 };
    if (rc)
      return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MICM_HIP_OK;
  }

  int micm_hip_adjoint(const micm_hip_mechanism* m, micm_hip_state* s, const micm_hip_rosenbrock_parameters* params,
                       const micm_hip_tape* tape, const double* seed, double* grad_variables,
                       double* grad_log_rate_constants)
  {
    if (int rc = check_mech(m))
      return rc;
    if (!s || s->mech != m || !params || !tape || tape->mech != m || tape->n_cells != s->n_cells || !seed ||
        !grad_variables || !grad_log_rate_constants)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "bad state, parameters, tape or buffers");
    if (params->stages < 1 || params->stages > 6)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "stages must be 1..6");
    if (!m->t.coop_levels_ok)
      return fail(MICM_HIP_ERR_UNSUPPORTED, "the adjoint needs the level schedules (positions above 16 bits)");
    const DevMech& d = m->dev;
    const DevCoop& c = m->coop;
    const size_t lds = adjoint_lds_doubles(c.nnz_alu, d.n_species, static_cast<int>(params->stages), d.n_reactions,
                                           c.n_terms) *
                       sizeof(double);
    if (lds > lds_limit())
      return fail(MICM_HIP_ERR_UNSUPPORTED, "the adjoint's per-cell working set (" + std::to_string(lds) +
                                                " bytes) does not fit one workgroup's LDS");
    const DevParams p = make_params(*params, 1.0, s->rtol);
    constexpr int kAdjThreads = 256;
    const void* fn = reinterpret_cast<const void*>(k_adjoint<kAdjThreads>);
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    int dev = 0, cus = 0, per_cu = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adjoint<kAdjThreads>, kAdjThreads, lds));
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(s->n_cells, static_cast<int64_t>(cus) *
                                                                                   std::max(per_cu, 1) * 4));
    const Tape tp{ tape->records, tape->counts, tape->max_records };
    hipLaunchKernelGGL(k_adjoint<kAdjThreads>, dim3(static_cast<unsigned>(blocks)), dim3(kAdjThreads), lds, s->stream,
                       d, c, m->adj, p, s->n_cells, s->k, tp, seed, grad_variables, grad_log_rate_constants);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MICM_HIP_OK;
  }

  // device-to-device transfers of grouped arrays (the state's stream, then it waits)
  int micm_hip_state_set_variables_device(micm_hip_state* s, const double* src)
  {
    if (!s || !src)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipMemcpyAsync(s->y, src, sizeof(double) * grouped_doubles(s, s->mech->dev.n_species),
                           hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MICM_HIP_OK;
  }
  int micm_hip_state_get_variables_device(const micm_hip_state* s, double* dst)
  {
    if (!s || !dst)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipMemcpyAsync(dst, s->y, sizeof(double) * grouped_doubles(s, s->mech->dev.n_species),
                           hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MICM_HIP_OK;
  }
  int micm_hip_state_set_rate_constants_device(micm_hip_state* s, const double* src)
  {
    if (!s || !src)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipMemcpyAsync(s->k, src, sizeof(double) * grouped_doubles(s, s->mech->dev.n_reactions),
                           hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MICM_HIP_OK;
  }
  int micm_hip_state_get_rate_constants_device(const micm_hip_state* s, double* dst)
  {
    if (!s || !dst)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipMemcpyAsync(dst, s->k, sizeof(double) * grouped_doubles(s, s->mech->dev.n_reactions),
                           hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MICM_HIP_OK;
  }

  int micm_hip_packed_create(int32_t n_states, micm_hip_state* const* states, int32_t form, micm_hip_packed** out)
  {
    if (!out || n_states < 1 || !states)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "need at least one state");
    if (form != MICM_HIP_SOLVE_PERSISTENT && form != MICM_HIP_SOLVE_COOPERATIVE)
      return fail(MICM_HIP_ERR_UNSUPPORTED, "packed solves use the persistent or the cooperative form");
    for (int32_t i = 0; i < n_states; ++i)
      if (!states[i])
        return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null state");
    auto p = std::make_unique<micm_hip_packed>();
    p->states.assign(states, states + n_states);
    p->form = form;
    HIP_TRY(hipStreamCreate(&p->own_stream));
    p->origin = p->own_stream;
    *out = p.release();
    return MICM_HIP_OK;
  }

  int micm_hip_packed_solve(micm_hip_packed* pk, double time_step, const micm_hip_rosenbrock_parameters* params)
  {
    if (!pk || !params)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "bad packed object or parameters");
    if (params->stages < 1 || params->stages > 6)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "stages must be 1..6");
    std::vector<int64_t> shapes;
    for (micm_hip_state* s : pk->states)
    {
      if (int rc = prepare_solve(s, pk->form, static_cast<int>(params->stages)))
        return rc;
      // everything the captured launches bake in: launch shapes, the relative tolerance
      // (a kernel argument) and the scratch buffers prepare_solve may have reallocated
      int64_t rtol_bits = 0;
      std::memcpy(&rtol_bits, &s->rtol, sizeof rtol_bits);
      shapes.push_back(s->handoff_steps);
      shapes.push_back(s->coop_threads);
      shapes.push_back(s->coop_stage);
      shapes.push_back(s->lu_blocking);
      shapes.push_back(static_cast<int64_t>(s->n_slots));
      shapes.push_back(rtol_bits);
      shapes.push_back(static_cast<int64_t>(reinterpret_cast<intptr_t>(s->coop_k)));
      shapes.push_back(s->coop_k_rows);
      shapes.push_back(static_cast<int64_t>(reinterpret_cast<intptr_t>(s->slot_blob)));
      // the first Solve of a state runs its cells in natural order, the later ones in the
      // previous Solve's cost order: the graph captured first is replaced once the order exists
      shapes.push_back(s->has_order ? 1 : 0);
    }
    const bool same = pk->captured && pk->dt == time_step && pk->shapes == shapes &&
                      std::memcmp(&pk->params, params, sizeof *params) == 0;
    if (!same)
    {
      if (pk->exec)
        (void)hipGraphExecDestroy(pk->exec);
      if (pk->graph)
        (void)hipGraphDestroy(pk->graph);
      pk->exec = nullptr;
      pk->graph = nullptr;
      pk->captured = false;
      HIP_TRY(hipStreamBeginCapture(pk->origin, hipStreamCaptureModeThreadLocal));
      int rc = MICM_HIP_OK;
      for (size_t i = 0; rc == MICM_HIP_OK && i < pk->states.size(); ++i)
      {
        micm_hip_state* s = pk->states[i];
        const DevParams p = make_params(*params, time_step, s->rtol);
        rc = enqueue_solve(s, p, pk->form, false, pk->origin);
      }
      hipGraph_t graph = nullptr;
      const hipError_t end = hipStreamEndCapture(pk->origin, &graph);
      if (rc != MICM_HIP_OK)
      {
        if (graph)
          (void)hipGraphDestroy(graph);
        return rc;
      }
      if (end != hipSuccess || !graph)
        return fail(MICM_HIP_ERR_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(end));
      pk->graph = graph;
      HIP_TRY(hipGraphInstantiate(&pk->exec, pk->graph, nullptr, nullptr, 0));
      pk->captured = true;
      pk->dt = time_step;
      pk->params = *params;
      pk->shapes = shapes;
    }
    pk->replayed = same;
    HIP_TRY(hipGraphLaunch(pk->exec, pk->origin));
    HIP_TRY(hipStreamSynchronize(pk->origin));
    return MICM_HIP_OK;
  }

  int micm_hip_packed_set_stream(micm_hip_packed* pk, void* stream)
  {
    if (!pk || !stream)
      return fail(MICM_HIP_ERR_INVALID_ARGUMENT, "null packed object or stream (a graph cannot be captured on the "
                                                 "legacy default stream)");
    pk->origin = static_cast<hipStream_t>(stream);
    return MICM_HIP_OK;
  }

  int micm_hip_packed_replayed(const micm_hip_packed* pk)
  {
    return pk && pk->replayed ? 1 : 0;
  }

  void micm_hip_packed_destroy(micm_hip_packed* pk)
  {
    if (!pk)
      return;
    if (pk->exec)
      (void)hipGraphExecDestroy(pk->exec);
    if (pk->graph)
      (void)hipGraphDestroy(pk->graph);
    if (pk->own_stream)
    {
      (void)hipStreamSynchronize(pk->own_stream);
      (void)hipStreamDestroy(pk->own_stream);
    }
    delete pk;
  }

}
