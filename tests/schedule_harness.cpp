// schedule_harness.cpp -- CPU execution of the cooperative form's level schedules
// (micm_amd/csrc/mechanism.cpp), for tests/test_schedule.py.
//
// Every level reads the values as they stood when the level began, so a task that
// depended on another task of its own level would read a stale value and the result
// would differ from the sequential Doolittle factorisation.
#include "../micm_amd/csrc/mechanism.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

using micm_hip::LevelSchedule;
using micm_hip::MechanismTables;

namespace
{
  // out = (out - sum A[lo] X[hi]) / A[diag], every task of a level from a snapshot
  void run(const LevelSchedule& s, const double* A_now, std::vector<double>& A_state, std::vector<double>& X_state,
           bool out_is_a, bool x_is_a)
  {
    for (int32_t l = 0; l < s.n_levels(); ++l)
    {
      const std::vector<double> A0 = A_state, X0 = X_state;
      const double* A = out_is_a ? A0.data() : A_now;
      const double* X = x_is_a ? A0.data() : X0.data();
      for (int32_t t = s.offsets[l]; t < s.offsets[l + 1]; ++t)
      {
        const uint32_t o = s.tasks[2 * t] & 0xffffu, d = s.tasks[2 * t] >> 16;
        double v = out_is_a ? A0[o] : X0[o];
        for (uint32_t q = s.tasks[2 * t + 1]; q < s.tasks[2 * t + 3]; ++q)
          v -= A[s.terms[q] & 0xffffu] * X[s.terms[q] >> 16];
        if (d < micm_hip::kUnitDivisor)
          v = v / A[d];
        (out_is_a ? A_state : X_state)[o] = v;
      }
    }
  }
}  // namespace

extern "C"
{
  // sizes: n_species, nnz_jac, nnz_alu, nnz_lower, nnz_upper, levels (lu, fwd, bwd), ok
  int harness_sizes(const micm_hip_mechanism_desc* desc, int64_t* sizes)
  {
    const MechanismTables t = micm_hip::compile_mechanism(*desc);
    sizes[0] = t.n_species;
    sizes[1] = t.nnz_jacobian();
    sizes[2] = t.nnz_alu();
    sizes[3] = t.nnz_lower();
    sizes[4] = t.nnz_upper();
    sizes[5] = t.lv_lu.n_levels();
    sizes[6] = t.lv_fwd.n_levels();
    sizes[7] = t.lv_bwd.n_levels();
    sizes[8] = t.coop_levels_ok ? 1 : 0;
    return 0;
  }

  // One cell: -J (state order, CSR) with alpha added on the diagonal, factorised in
  // place by the LU schedule; L (unit diagonal) and U scattered out in the reference's
  // storage order; then b is solved in place by the substitution schedules.
  int harness_factor_solve(const micm_hip_mechanism_desc* desc, const double* jac, double alpha, double* lower,
                           double* upper, double* b)
  {
    const MechanismTables t = micm_hip::compile_mechanism(*desc);
    const int32_t n = static_cast<int32_t>(t.n_species);
    std::vector<double> alu(t.nnz_alu(), 0.0), none;
    for (size_t e = 0; e < t.jac_to_alu.size(); ++e)
      alu[t.jac_to_alu[e]] = jac[e];
    for (int32_t i = 0; i < n; ++i)
      alu[t.alu_diag[i]] += alpha;  // as AlphaMinusJacobian on -J
    run(t.lv_lu, nullptr, alu, none, true, true);
    for (size_t p = 0; p < t.alu_lower.size(); ++p)
      lower[p] = t.alu_lower[p] >= 0 ? alu[t.alu_lower[p]] : 1.0;
    for (size_t p = 0; p < t.alu_upper.size(); ++p)
      upper[p] = alu[t.alu_upper[p]];
    std::vector<double> x(b, b + n);
    run(t.lv_fwd, alu.data(), none, x, false, false);
    run(t.lv_bwd, alu.data(), none, x, false, false);
    std::memcpy(b, x.data(), sizeof(double) * n);
    return 0;
  }
}

extern "C"
{
  // The factorisation of harness_factor_solve by the LU schedule of blocking `b`
  // (index into kLuBlockings); returns the number of levels, or -1 for a bad index.
  int harness_factor_blocked(const micm_hip_mechanism_desc* desc, int32_t b, const double* jac, double alpha,
                             double* lower, double* upper)
  {
    const MechanismTables t = micm_hip::compile_mechanism(*desc);
    if (b < 0 || static_cast<size_t>(b) >= t.lv_lu_blocked.size())
      return -1;
    const int32_t n = static_cast<int32_t>(t.n_species);
    std::vector<double> alu(t.nnz_alu(), 0.0), none;
    for (size_t e = 0; e < t.jac_to_alu.size(); ++e)
      alu[t.jac_to_alu[e]] = jac[e];
    for (int32_t i = 0; i < n; ++i)
      alu[t.alu_diag[i]] += alpha;
    run(t.lv_lu_blocked[b], nullptr, alu, none, true, true);
    for (size_t p = 0; p < t.alu_lower.size(); ++p)
      lower[p] = t.alu_lower[p] >= 0 ? alu[t.alu_lower[p]] : 1.0;
    for (size_t p = 0; p < t.alu_upper.size(); ++p)
      upper[p] = alu[t.alu_upper[p]];
    return t.lv_lu_blocked[b].n_levels();
  }
}

namespace
{
  // True when a task of some level writes a position that another task of the same level
  // reads or writes (a race on the device, invisible to run()'s level snapshots).
  bool same_level_conflict(const LevelSchedule& s)
  {
    for (int32_t l = 0; l < s.n_levels(); ++l)
    {
      std::vector<std::pair<uint32_t, int32_t>> written;  // position, task
      for (int32_t t = s.offsets[l]; t < s.offsets[l + 1]; ++t)
        written.emplace_back(s.tasks[2 * t] & 0xffffu, t);
      std::sort(written.begin(), written.end());
      for (size_t i = 1; i < written.size(); ++i)
        if (written[i].first == written[i - 1].first)
          return true;
      auto writer = [&](uint32_t pos) {
        auto it = std::lower_bound(written.begin(), written.end(), std::make_pair(pos, INT32_MIN));
        return it != written.end() && it->first == pos ? it->second : -1;
      };
      for (int32_t t = s.offsets[l]; t < s.offsets[l + 1]; ++t)
      {
        const uint32_t d = s.tasks[2 * t] >> 16;
        if (d < micm_hip::kUnitDivisor && writer(d) >= 0 && writer(d) != t)
          return true;
        for (uint32_t q = s.tasks[2 * t + 1]; q < s.tasks[2 * t + 3]; ++q)
          for (uint32_t pos : { s.terms[q] & 0xffffu, s.terms[q] >> 16 })
            if (writer(pos) >= 0 && writer(pos) != t)
              return true;
      }
    }
    return false;
  }
}  // namespace

extern "C"
{
  // The factorisation of harness_factor_solve by every LU schedule with capped levels,
  // blocking b and cap c at slot s = b * kNumLuCaps + c of lower / upper (nnz_lower /
  // nnz_upper values each) and levels: the number of levels, 0 where there is no capped
  // schedule (no level exceeds the cap), -2 where a level holds more tasks than the cap or
  // two tasks of one level touch one position.  Returns the number of slots.
  int harness_factor_capped(const micm_hip_mechanism_desc* desc, const double* jac, double alpha, double* lower,
                            double* upper, int32_t* levels)
  {
    const MechanismTables t = micm_hip::compile_mechanism(*desc);
    const int32_t n = static_cast<int32_t>(t.n_species);
    const size_t slots = t.lv_lu_capped.size();
    for (size_t k = 0; k < slots; ++k)
    {
      const LevelSchedule& s = t.lv_lu_capped[k];
      const int cap = micm_hip::kLuCaps[k % micm_hip::kNumLuCaps];
      levels[k] = s.n_levels();
      if (s.n_levels() == 0)
        continue;
      for (int32_t l = 0; l < s.n_levels(); ++l)
        if (s.offsets[l + 1] - s.offsets[l] > cap)
          levels[k] = -2;
      if (same_level_conflict(s))
        levels[k] = -2;
      std::vector<double> alu(t.nnz_alu(), 0.0), none;
      for (size_t e = 0; e < t.jac_to_alu.size(); ++e)
        alu[t.jac_to_alu[e]] = jac[e];
      for (int32_t i = 0; i < n; ++i)
        alu[t.alu_diag[i]] += alpha;
      run(s, nullptr, alu, none, true, true);
      double* lo = lower + k * t.alu_lower.size();
      double* up = upper + k * t.alu_upper.size();
      for (size_t p = 0; p < t.alu_lower.size(); ++p)
        lo[p] = t.alu_lower[p] >= 0 ? alu[t.alu_lower[p]] : 1.0;
      for (size_t p = 0; p < t.alu_upper.size(); ++p)
        up[p] = alu[t.alu_upper[p]];
    }
    return static_cast<int>(slots);
  }

  // 1 when some level of the (uncapped) LU schedule of blocking `b` has two tasks
  // touching one position, 0 when none has, -1 for a bad index.
  int harness_blocked_conflict(const micm_hip_mechanism_desc* desc, int32_t b)
  {
    const MechanismTables t = micm_hip::compile_mechanism(*desc);
    if (b < 0 || static_cast<size_t>(b) >= t.lv_lu_blocked.size())
      return -1;
    return same_level_conflict(t.lv_lu_blocked[b]) ? 1 : 0;
  }
}

namespace
{
  // The forward substitution as the device runs it with a column sweep: the levels before
  // the sweep, then the sweep lane by lane and column by column (micm_amd/csrc/coop.hpp,
  // coop_sweep).
  void forward_swept(const MechanismTables& t, const std::vector<double>& alu, std::vector<double>& x)
  {
    const micm_hip::ColumnSweep& sw = t.fwd_sweep;
    LevelSchedule head = t.lv_fwd;
    head.offsets.resize(static_cast<size_t>(sw.first_level) + 1);
    std::vector<double> none;
    run(head, alu.data(), none, x, false, false);
    const int lanes = micm_hip::kSweepLanes;
    std::vector<double> v(lanes, 0.0), xv(lanes, 0.0);
    for (int r = 0; r < lanes; ++r)
    {
      const uint32_t row = sw.lanes[4 * r] & 0xffffu, src = sw.lanes[4 * r + 3];
      v[r] = row != 0xffffu ? x[row] : 0.0;
      xv[r] = r < sw.n_cols && !(src & micm_hip::kSweepFromLane) ? x[src] : 0.0;
    }
    for (int c = 0; c < sw.n_cols; ++c)
    {
      const uint32_t s = sw.lanes[4 * c + 3];
      const double xc = (s & micm_hip::kSweepFromLane) ? v[s & 63u] : xv[c];
      for (int r = 0; r < lanes; ++r)
      {
        const uint64_t mask = sw.lanes[4 * r + 1] | (static_cast<uint64_t>(sw.lanes[4 * r + 2]) << 32);
        if (!((mask >> c) & 1u))
          continue;
        const uint32_t start = sw.lanes[4 * r] >> 16;
        const uint64_t below = mask & ((uint64_t(1) << c) - 1u);
        v[r] = v[r] - alu[start + static_cast<uint32_t>(__builtin_popcountll(below))] * xc;
      }
    }
    for (int r = 0; r < lanes; ++r)
      if ((sw.lanes[4 * r] & 0xffffu) != 0xffffu)
        x[sw.lanes[4 * r] & 0xffffu] = v[r];
  }
}  // namespace

extern "C"
{
  // harness_factor_solve with the forward substitution's column sweep and the back
  // substitution's task chain; returns the first swept level (-1: no sweep, and b
  // untouched).
  int harness_solve_swept(const micm_hip_mechanism_desc* desc, const double* jac, double alpha, double* b)
  {
    const MechanismTables t = micm_hip::compile_mechanism(*desc);
    if (t.fwd_sweep.first_level < 0)
      return -1;
    const int32_t n = static_cast<int32_t>(t.n_species);
    std::vector<double> alu(t.nnz_alu(), 0.0), none;
    for (size_t e = 0; e < t.jac_to_alu.size(); ++e)
      alu[t.jac_to_alu[e]] = jac[e];
    for (int32_t i = 0; i < n; ++i)
      alu[t.alu_diag[i]] += alpha;
    run(t.lv_lu, nullptr, alu, none, true, true);
    alu.push_back(0.0);  // the zero one past the factor
    std::vector<double> x(b, b + n);
    forward_swept(t, alu, x);
    // the back substitution as the device runs it: the task chain in its order, each
    // task's late product formed from the previous task's result, then the other levels
    const micm_hip::TaskChain& ch = t.bwd_chain;
    double x_prev = 0.0;
    for (size_t q = 0; q < ch.lane_terms.size() / micm_hip::kChainLanes; ++q)
    {
      const uint32_t o = ch.headers[2 * q] & 0xffffu, d = ch.headers[2 * q] >> 16,
                     nt = ch.headers[2 * q + 1] & 0xffffu, late = ch.headers[2 * q + 1] >> 16;
      double v = x[o];
      for (uint32_t k = 0; k < nt; ++k)
      {
        const uint32_t tr = ch.lane_terms[q * micm_hip::kChainLanes + k];
        v -= alu[tr & 0xffffu] * (k == late ? x_prev : x[tr >> 16]);
      }
      if (d < micm_hip::kUnitDivisor)
        v = v / alu[d];
      x[o] = v;
      x_prev = v;
    }
    LevelSchedule tail = t.lv_bwd;
    tail.offsets.erase(tail.offsets.begin(), tail.offsets.begin() + ch.levels);
    run(tail, alu.data(), none, x, false, false);
    std::memcpy(b, x.data(), sizeof(double) * n);
    return t.fwd_sweep.first_level;
  }

  // The back substitution's task chain against its level schedule: the chain holds the
  // tasks of its levels in level order with their terms; a task's late term is the one
  // reading the previous task's output, and every other term reads a value whose last
  // writer comes at least two tasks earlier (its product is formed a task ahead).
  // Returns 0, or 1 + the index of the first task that breaks this.
  int harness_chain_check(const micm_hip_mechanism_desc* desc)
  {
    const MechanismTables t = micm_hip::compile_mechanism(*desc);
    const micm_hip::TaskChain& ch = t.bwd_chain;
    const LevelSchedule& s = t.lv_bwd;
    if (ch.levels == 0)
      return ch.headers.empty() && ch.lane_terms.empty() ? 0 : 1;
    const size_t nt = ch.lane_terms.size() / micm_hip::kChainLanes;
    if (nt != static_cast<size_t>(s.offsets[ch.levels]) || nt > static_cast<size_t>(micm_hip::kChainTasks) ||
        ch.headers.size() != 2 * static_cast<size_t>(micm_hip::kChainTasks))
      return 1;
    std::vector<int64_t> last_writer(t.n_species, -1);
    for (size_t q = 0; q < nt; ++q)
    {
      const uint32_t first = s.tasks[2 * q + 1], n = s.tasks[2 * q + 3] - first;
      if (ch.headers[2 * q] != s.tasks[2 * q] || (ch.headers[2 * q + 1] & 0xffffu) != n || n > micm_hip::kChainLanes)
        return static_cast<int>(q) + 1;
      const uint32_t prev_out = q > 0 ? (ch.headers[2 * (q - 1)] & 0xffffu) : 0xffffffffu;
      for (uint32_t k = 0; k < n; ++k)
      {
        const uint32_t tr = ch.lane_terms[q * micm_hip::kChainLanes + k], col = tr >> 16;
        if (tr != s.terms[first + k])
          return static_cast<int>(q) + 1;
        const bool late = col == prev_out;
        if (late != ((ch.headers[2 * q + 1] >> 16) == k))
          return static_cast<int>(q) + 1;
        if (!late && last_writer[col] >= 0 && last_writer[col] + 2 > static_cast<int64_t>(q))
          return static_cast<int>(q) + 1;
      }
      last_writer[ch.headers[2 * q] & 0xffffu] = static_cast<int64_t>(q);
    }
    return 0;
  }

  // Per level of schedule `which` (0 LU, 1 forward, 2 backward): tasks and the
  // largest number of terms of one task; returns the number of levels written.
  int harness_level_stats(const micm_hip_mechanism_desc* desc, int32_t which, int64_t* tasks, int64_t* max_terms,
                          int64_t capacity)
  {
    const MechanismTables t = micm_hip::compile_mechanism(*desc);
    const LevelSchedule& s = which == 0 ? t.lv_lu : (which == 1 ? t.lv_fwd : t.lv_bwd);
    const int32_t nl = s.n_levels();
    for (int32_t l = 0; l < nl && l < capacity; ++l)
    {
      tasks[l] = s.offsets[l + 1] - s.offsets[l];
      int64_t m = 0;
      for (int32_t q = s.offsets[l]; q < s.offsets[l + 1]; ++q)
        m = std::max<int64_t>(m, s.tasks[2 * q + 3] - s.tasks[2 * q + 1]);
      max_terms[l] = m;
    }
    return nl;
  }
}

extern "C"
{
  // Longest dependency chain of schedule `which`, weighting each task by its number
  // of terms plus `base`: the lower bound any level assignment can reach.
  double harness_critical_path(const micm_hip_mechanism_desc* desc, int32_t which, double base)
  {
    const MechanismTables t = micm_hip::compile_mechanism(*desc);
    const LevelSchedule& s = which == 0 ? t.lv_lu : (which == 1 ? t.lv_fwd : t.lv_bwd);
    const size_t n_pos = which == 0 ? t.alu_col.size() : static_cast<size_t>(t.n_species);
    std::vector<double> finish(n_pos, 0.0);
    double best = 0.0;
    for (int32_t l = 0; l < s.n_levels(); ++l)
      for (int32_t q = s.offsets[l]; q < s.offsets[l + 1]; ++q)
      {
        const uint32_t o = s.tasks[2 * q] & 0xffffu, d = s.tasks[2 * q] >> 16;
        double start = finish[o];  // an earlier segment of the same sum
        for (uint32_t k = s.tasks[2 * q + 1]; k < s.tasks[2 * q + 3]; ++k)
        {
          if (which == 0)
            start = std::max(start, std::max(finish[s.terms[k] & 0xffffu], finish[s.terms[k] >> 16]));
          else
            start = std::max(start, finish[s.terms[k] >> 16]);
        }
        if (which == 0 && d < micm_hip::kUnitDivisor)
          start = std::max(start, finish[d]);
        finish[o] = start + base + (s.tasks[2 * q + 3] - s.tasks[2 * q + 1]);
        best = std::max(best, finish[o]);
      }
    return best;
  }
}
