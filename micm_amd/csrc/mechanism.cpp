// mechanism.cpp -- symbolic compilation of a mechanism (host side).
//
//   reaction sort by rate-constant type      include/micm/solver/solver_builder.inl:274-329
//   species map, diagonal Markowitz order    solver_builder.inl:78-119, include/micm/solver/linear_solver.inl:29-124
//   ProcessSet reactant / product tables     include/micm/process/process_set.hpp:256-422
//   Jacobian non-zeros and flat ids          process_set.hpp:425-492, include/micm/util/jacobian.hpp:15-32
//   Doolittle fill-in and schedule           include/micm/solver/lu_decomposition_doolittle.inl:69-299
//   triangular-solve tables                  linear_solver.inl:166-224
//   rate-constant store                      include/micm/process/reaction_rate_store.hpp:221-368
#include "mechanism.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <queue>
#include <stdexcept>

namespace micm_hip
{
  namespace
  {
    constexpr double kBoltzmann = 1.380649e-23;
    constexpr double kAvogadro = 6.02214076e23;
    constexpr double kGasConstant = kBoltzmann * kAvogadro;  // include/micm/util/constants.hpp:9-11
    constexpr double kPi = 3.14159265358979323846;

    bool has(const std::vector<int32_t>& sorted, int32_t x)
    {
      return std::binary_search(sorted.begin(), sorted.end(), x);
    }

    void insert_sorted(std::vector<int32_t>& v, int32_t x)
    {
      auto it = std::lower_bound(v.begin(), v.end(), x);
      if (it == v.end() || *it != x)
        v.insert(it, x);
    }

    void erase_sorted(std::vector<int32_t>& v, int32_t x)
    {
      auto it = std::lower_bound(v.begin(), v.end(), x);
      if (it != v.end() && *it == x)
        v.erase(it);
    }

    // Minimum-Markowitz-cost pivot order on the directed graph of the non-zero
    // pattern (edge v->c iff entry (v, c) is non-zero).  The cost
    // (row_degree - 1) * (column_degree - 1) is evaluated in unsigned 64-bit
    // arithmetic and ties go to the lowest index, as in the reference.
    // Returns perm[new] = old.
    std::vector<int64_t> markowitz_order(int32_t n, const std::vector<std::pair<int32_t, int32_t>>& edges)
    {
      std::vector<std::vector<int32_t>> out(n), in(n);
      for (auto [r, c] : edges)
      {
        insert_sorted(out[r], c);
        insert_sorted(in[c], r);
      }
      std::vector<uint64_t> rdeg(n), cdeg(n);
      for (int32_t v = 0; v < n; ++v)
      {
        rdeg[v] = out[v].size();
        cdeg[v] = in[v].size();
      }
      std::vector<char> alive(n, 1);
      std::vector<int64_t> perm;
      perm.reserve(n);
      std::vector<int32_t> cols, rows;
      for (int32_t step = 0; step < n; ++step)
      {
        int32_t pivot = -1;
        uint64_t best = 0;
        for (int32_t v = 0; v < n; ++v)
        {
          if (!alive[v])
            continue;
          const uint64_t cost = (rdeg[v] - 1u) * (cdeg[v] - 1u);
          if (pivot < 0 || cost < best)
          {
            best = cost;
            pivot = v;
          }
        }
        perm.push_back(pivot);
        alive[pivot] = 0;
        cols.clear();
        rows.clear();
        for (int32_t c : out[pivot])
          if (c != pivot && alive[c])
            cols.push_back(c);
        for (int32_t r : in[pivot])
          if (r != pivot && alive[r])
            rows.push_back(r);
        for (int32_t r : rows)
          for (int32_t c : cols)
            if (!has(out[r], c))
            {
              insert_sorted(out[r], c);
              insert_sorted(in[c], r);
              ++rdeg[r];
              ++cdeg[c];
            }
        for (int32_t c : cols)
        {
          erase_sorted(in[c], pivot);
          --cdeg[c];
        }
        for (int32_t r : rows)
        {
          erase_sorted(out[r], pivot);
          --rdeg[r];
        }
      }
      return perm;
    }

    // CSR position of (row, col) in a sorted-row pattern.
    int32_t position(const std::vector<int32_t>& row_ptr, const std::vector<int32_t>& col, int32_t r, int32_t c)
    {
      auto b = col.begin() + row_ptr[r], e = col.begin() + row_ptr[r + 1];
      auto it = std::lower_bound(b, e, c);
      if (it == e || *it != c)
        return -1;
      return static_cast<int32_t>(it - col.begin());
    }

    void to_csr(const std::vector<std::vector<int32_t>>& rows, std::vector<int32_t>& ptr, std::vector<int32_t>& col)
    {
      ptr.assign(1, 0);
      col.clear();
      for (const auto& r : rows)
      {
        col.insert(col.end(), r.begin(), r.end());
        ptr.push_back(static_cast<int32_t>(col.size()));
      }
    }
  }  // namespace

  namespace
  {
    struct LevelTask
    {
      int32_t level = 0;
      uint32_t out = 0, diag = 0;
      std::vector<uint32_t> terms;
    };

    // Tasks grouped by level; within a level the longest first, so that a lane's
    // round-robin share is balanced.  tasks[2 t] = out | diag << 16, tasks[2 t + 1] =
    // first term; a sentinel task closes the last term range.
    // `level` (one per task) replaces the tasks' own levels when given; `pad` fills the
    // record slots past a task's last term.
    LevelSchedule pack_levels(const std::vector<LevelTask>& v, const std::vector<int32_t>* level = nullptr,
                              uint32_t pad = 0)
    {
      auto level_of = [&](size_t i) { return level ? (*level)[i] : v[i].level; };
      std::vector<uint32_t> idx(v.size());
      for (size_t i = 0; i < v.size(); ++i)
        idx[i] = static_cast<uint32_t>(i);
      std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        if (level_of(a) != level_of(b))
          return level_of(a) < level_of(b);
        return v[a].terms.size() > v[b].terms.size();
      });
      LevelSchedule s;
      s.offsets.push_back(0);
      int32_t lv = 0;
      for (size_t q = 0; q < v.size(); ++q)
      {
        const LevelTask& x = v[idx[q]];
        while (level_of(idx[q]) > lv)
        {
          s.offsets.push_back(static_cast<int32_t>(q));
          ++lv;
        }
        s.tasks.push_back(x.out | (x.diag << 16));
        s.tasks.push_back(static_cast<uint32_t>(s.terms.size()));
        s.terms.insert(s.terms.end(), x.terms.begin(), x.terms.end());
        s.max_terms = std::max<int32_t>(s.max_terms, static_cast<int32_t>(x.terms.size()));
      }
      if (!v.empty())
        s.offsets.push_back(static_cast<int32_t>(v.size()));
      s.tasks.push_back(0xffffffffu);
      s.tasks.push_back(static_cast<uint32_t>(s.terms.size()));
      for (size_t l = 0; l + 1 < s.offsets.size(); ++l)
        s.max_level_terms = std::max<int32_t>(
            s.max_level_terms, static_cast<int32_t>(s.tasks[2 * s.offsets[l + 1] + 1] - s.tasks[2 * s.offsets[l] + 1]));
      // records: the task's header and its first terms in one 64-byte read
      s.records.clear();
      for (size_t q = 0; q + 1 < s.tasks.size() / 2; ++q)
      {
        const uint32_t first = s.tasks[2 * q + 1], count = s.tasks[2 * q + 3] - first;
        s.records.push_back(s.tasks[2 * q]);
        s.records.push_back(first);
        s.records.push_back(count);
        for (uint32_t r = 0; r < static_cast<uint32_t>(kInlineTerms); ++r)
          s.records.push_back(r < count ? s.terms[first + r] : pad);
      }
      // the sentinel: no output, no terms
      s.records.push_back(0u | (kNoDivisor << 16));
      s.records.push_back(static_cast<uint32_t>(s.terms.size()));
      s.records.push_back(0u);
      for (int r = 0; r < kInlineTerms; ++r)
        s.records.push_back(0u);
      // sixteen entries naming position 0 let the batches of eight read past a task's last term
      s.terms.insert(s.terms.end(), 16, 0u);
      return s;
    }

    // Splits one element's ordered products into segments by the level at which
    // their operands become readable: term q runs at level max(level of term q - 1,
    // ready[q]).  Each segment is a task that continues the element's partial sum,
    // kept in the element's own position between levels, so every subtraction happens
    // in the same order and with the same rounding as in one long task, but the sum
    // starts as soon as its first operands exist instead of when its last one does.
    // The last segment runs no earlier than `final_min` and carries the divisor; an
    // element without products gets one task that only divides.  Returns the level
    // of the last segment.
    // Shortest segment (except an element's last) of the substitution sums; the LU's is a
    // choice (kLuBlockings): its many short segments can cost more in per-task overhead
    // than they save in chain length, the substitutions' do not (measured on MI355X for
    // 60/200, TS1 and 210/560).
    constexpr size_t kSolveSegmentMin = 1;

    int32_t add_segments(std::vector<LevelTask>& out, uint32_t pos, uint32_t diag, const std::vector<uint32_t>& terms,
                         const std::vector<int32_t>& ready, int32_t final_min, size_t min_len)
    {
      const size_t first = out.size();
      int32_t lv = 0;
      for (size_t q = 0; q < terms.size();)
      {
        lv = std::max(lv, ready[q]);
        size_t e = q + 1;
        while (e < terms.size() && (ready[e] <= lv || e - q < min_len))
        {
          lv = std::max(lv, ready[e]);
          ++e;
        }
        LevelTask seg;
        seg.out = pos;
        seg.diag = kNoDivisor;
        seg.level = lv;
        seg.terms.assign(terms.begin() + static_cast<std::ptrdiff_t>(q), terms.begin() + static_cast<std::ptrdiff_t>(e));
        out.push_back(std::move(seg));
        q = e;
      }
      if (out.size() == first)
      {
        LevelTask seg;
        seg.out = pos;
        seg.level = 0;
        out.push_back(std::move(seg));
      }
      LevelTask& last = out.back();
      last.level = std::max(last.level, final_min);
      last.diag = diag;
      return last.level;
    }

    // Order constraints of a schedule whose outputs, divisors and both operands of every
    // term index one array (the in-place factorisation): a task comes after the writer of
    // every value it reads (the value it read in the original schedule), after the
    // previous writer of its output, and after every reader of the value its output
    // replaces.  succ in CSR form; height = the longest path (in tasks) to the end.
    struct TaskOrder
    {
      std::vector<uint32_t> succ_ptr, succ, n_deps;
      std::vector<int32_t> height;
      size_t widest = 0;  // most tasks of one level of the original schedule
    };

    TaskOrder task_order(const std::vector<LevelTask>& v)
    {
      TaskOrder o;
      const size_t nt = v.size();
      int32_t n_levels = 0;
      for (const LevelTask& x : v)
        n_levels = std::max(n_levels, x.level + 1);
      std::vector<uint32_t> per_level(static_cast<size_t>(n_levels) + 1, 0);
      for (const LevelTask& x : v)
        ++per_level[static_cast<size_t>(x.level) + 1];
      for (uint32_t c : per_level)
        o.widest = std::max<size_t>(o.widest, c);
      for (size_t l = 1; l < per_level.size(); ++l)
        per_level[l] += per_level[l - 1];
      std::vector<uint32_t> order(nt);  // tasks in level order
      {
        std::vector<uint32_t> fill(per_level.begin(), per_level.end() - 1);
        for (size_t t = 0; t < nt; ++t)
          order[fill[static_cast<size_t>(v[t].level)]++] = static_cast<uint32_t>(t);
      }
      auto reads = [&](const LevelTask& x, auto&& f) {
        f(x.out);
        if (x.diag < kUnitDivisor)
          f(x.diag);
        for (uint32_t tr : x.terms)
        {
          f(tr & 0xffffu);
          f(tr >> 16);
        }
      };
      uint32_t n_pos = 0;
      for (const LevelTask& x : v)
        reads(x, [&](uint32_t i) { n_pos = std::max(n_pos, i + 1); });
      std::vector<int64_t> writer(n_pos, -1);
      std::vector<std::vector<uint32_t>> readers(n_pos);
      std::vector<std::pair<uint32_t, uint32_t>> edges;  // (before, after)
      std::vector<uint32_t> d;
      for (int32_t l = 0; l < n_levels; ++l)
      {
        const uint32_t b = per_level[static_cast<size_t>(l)], e = per_level[static_cast<size_t>(l) + 1];
        for (uint32_t q = b; q < e; ++q)
        {
          const uint32_t t = order[q];
          d.clear();
          reads(v[t], [&](uint32_t i) {
            if (writer[i] >= 0)
              d.push_back(static_cast<uint32_t>(writer[i]));
          });
          for (uint32_t r : readers[v[t].out])
            if (r != t)
              d.push_back(r);
          std::sort(d.begin(), d.end());
          d.erase(std::unique(d.begin(), d.end()), d.end());
          for (uint32_t x : d)
            edges.emplace_back(x, t);
        }
        for (uint32_t q = b; q < e; ++q)
          reads(v[order[q]], [&](uint32_t i) { readers[i].push_back(order[q]); });
        for (uint32_t q = b; q < e; ++q)
        {
          writer[v[order[q]].out] = order[q];
          readers[v[order[q]].out].clear();
        }
      }
      o.succ_ptr.assign(nt + 1, 0);
      o.n_deps.assign(nt, 0);
      for (const auto& ed : edges)
      {
        ++o.succ_ptr[ed.first + 1];
        ++o.n_deps[ed.second];
      }
      for (size_t t = 0; t < nt; ++t)
        o.succ_ptr[t + 1] += o.succ_ptr[t];
      o.succ.resize(edges.size());
      {
        std::vector<uint32_t> fill(o.succ_ptr.begin(), o.succ_ptr.end() - 1);
        for (const auto& ed : edges)
          o.succ[fill[ed.first]++] = ed.second;
      }
      o.height.assign(nt, 0);
      for (size_t q = nt; q-- > 0;)
      {
        const uint32_t t = order[q];
        for (uint32_t k = o.succ_ptr[t]; k < o.succ_ptr[t + 1]; ++k)
          o.height[t] = std::max(o.height[t], o.height[o.succ[k]] + 1);
      }
      return o;
    }

    // New levels for the tasks such that no level holds more than `cap`: a list schedule
    // that fills each level with the ready tasks on the longest remaining paths first.
    // The tasks and their terms are unchanged, so every result is the same bit for bit.
    std::vector<int32_t> cap_levels(const std::vector<LevelTask>& v, const TaskOrder& o, size_t cap)
    {
      const size_t nt = v.size();
      std::vector<int32_t> level(nt, 0);
      auto later = [&](uint32_t a, uint32_t b) {
        if (o.height[a] != o.height[b])
          return o.height[a] < o.height[b];
        if (v[a].level != v[b].level)
          return v[a].level > v[b].level;
        return a > b;
      };
      std::priority_queue<uint32_t, std::vector<uint32_t>, decltype(later)> ready(later);
      std::vector<uint32_t> pending(o.n_deps), next;
      for (size_t t = 0; t < nt; ++t)
        if (pending[t] == 0)
          ready.push(static_cast<uint32_t>(t));
      size_t placed = 0;
      for (int32_t lv = 0; placed < nt; ++lv)
      {
        next.clear();
        for (size_t k = 0; k < cap && !ready.empty(); ++k)
        {
          const uint32_t t = ready.top();
          ready.pop();
          level[t] = lv;
          ++placed;
          for (uint32_t q = o.succ_ptr[t]; q < o.succ_ptr[t + 1]; ++q)
            if (--pending[o.succ[q]] == 0)
              next.push_back(o.succ[q]);
        }
        for (uint32_t x : next)
          ready.push(x);
      }
      return level;
    }

    // The run of final levels of a forward substitution schedule that one wavefront
    // sweeps column by column (see ColumnSweep) with the largest gain: at most kSweepLanes
    // rows and 64 columns, every row's swept terms at consecutive ALU positions in
    // ascending columns, each below the row, and no divisions.  A level costs about as
    // much as eight columns of a sweep (a barrier and two LDS round trips against a few
    // instructions), so the run kept is the one with the most levels per column; none
    // when no run gains.
    ColumnSweep column_sweep(const LevelSchedule& s)
    {
      constexpr int32_t kColumnsPerLevel = 8;
      ColumnSweep best;
      int32_t best_gain = 0;
      const int32_t nl = s.n_levels();
      struct Row
      {
        uint32_t row;
        std::vector<uint32_t> terms;
      };
      for (int32_t first = nl - 1; first >= 0; --first)
      {
        // rows and their terms in level order
        std::vector<Row> rows;
        bool ok = true;
        for (int32_t l = first; l < nl && ok; ++l)
          for (int32_t q = s.offsets[l]; q < s.offsets[l + 1] && ok; ++q)
          {
            const uint32_t o = s.tasks[2 * q] & 0xffffu, d = s.tasks[2 * q] >> 16;
            if (d < kUnitDivisor)
              ok = false;
            Row* r = nullptr;
            for (Row& x : rows)
              if (x.row == o)
                r = &x;
            if (!r)
            {
              rows.push_back(Row{ o, {} });
              r = &rows.back();
            }
            r->terms.insert(r->terms.end(), s.terms.begin() + s.tasks[2 * q + 1], s.terms.begin() + s.tasks[2 * q + 3]);
          }
        if (!ok || rows.size() > static_cast<size_t>(kSweepLanes))
          break;
        std::vector<uint32_t> cols;
        for (const Row& r : rows)
          for (uint32_t tr : r.terms)
            cols.push_back(tr >> 16);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        if (cols.size() > 64)
          break;
        for (const Row& r : rows)
          for (size_t k = 0; k < r.terms.size() && ok; ++k)
          {
            const uint32_t pos = r.terms[k] & 0xffffu, col = r.terms[k] >> 16;
            ok = col < r.row && (k == 0 || (pos == (r.terms[0] & 0xffffu) + k && col > (r.terms[k - 1] >> 16)));
          }
        if (!ok)
          break;
        const int32_t gain = kColumnsPerLevel * (nl - first) - static_cast<int32_t>(cols.size());
        if (gain <= best_gain)
          continue;
        best_gain = gain;
        ColumnSweep sw;
        sw.first_level = first;
        sw.n_cols = static_cast<int32_t>(cols.size());
        sw.lanes.assign(4 * kSweepLanes, 0u);
        for (int r = 0; r < kSweepLanes; ++r)
          sw.lanes[4 * r] = 0xffffu;
        for (size_t r = 0; r < rows.size(); ++r)
        {
          uint64_t mask = 0;
          for (uint32_t tr : rows[r].terms)
            mask |= uint64_t(1) << (std::lower_bound(cols.begin(), cols.end(), tr >> 16) - cols.begin());
          const uint32_t start = rows[r].terms.empty() ? 0u : (rows[r].terms[0] & 0xffffu);
          sw.lanes[4 * r] = rows[r].row | (start << 16);
          sw.lanes[4 * r + 1] = static_cast<uint32_t>(mask);
          sw.lanes[4 * r + 2] = static_cast<uint32_t>(mask >> 32);
        }
        for (size_t c = 0; c < cols.size(); ++c)
        {
          uint32_t src = cols[c];
          for (size_t r = 0; r < rows.size(); ++r)
            if (rows[r].row == cols[c])
              src = static_cast<uint32_t>(r) | kSweepFromLane;
          sw.lanes[4 * c + 3] = src;
        }
        best = std::move(sw);
      }
      return best;
    }

    // The prefix of a back substitution schedule's levels that runs fastest as a task
    // chain (see TaskChain): a level costs about a barrier pair, three LDS round trips
    // and its instructions plus its longest sum, a chained task its instructions, its
    // division and its sum (about 1,700 and 1,000 clocks plus 8 a product, measured on
    // MI355X with 210/560), so the prefix kept is the one that saves the most by that
    // count: levels of one task gain, levels of several lose.  Tasks of at most
    // kChainLanes terms, at most kChainTasks tasks.
    TaskChain task_chain(const LevelSchedule& s)
    {
      constexpr int64_t kLevelCost = 1700, kTaskCost = 1000, kTermCost = 8;
      TaskChain best;
      const int32_t nl = s.n_levels();
      int64_t saved = 0, best_saved = 0;
      for (int32_t l = 0; l < nl && s.offsets[l + 1] <= kChainTasks; ++l)
      {
        int64_t longest = 0, tasks_cost = 0;
        bool ok = true;
        for (int32_t q = s.offsets[l]; q < s.offsets[l + 1]; ++q)
        {
          const int64_t n = s.tasks[2 * q + 3] - s.tasks[2 * q + 1];
          ok = ok && n <= kChainLanes;
          longest = std::max(longest, n);
          tasks_cost += kTaskCost + kTermCost * n;
        }
        if (!ok)
          break;
        saved += kLevelCost + kTermCost * longest - tasks_cost;
        if (saved > best_saved)
        {
          best_saved = saved;
          best.levels = l + 1;
        }
      }
      if (best.levels == 0)
        return best;
      uint32_t prev_out = 0xffffffffu;
      for (int32_t q = 0; q < s.offsets[best.levels]; ++q)
      {
        const uint32_t first = s.tasks[2 * q + 1], n = s.tasks[2 * q + 3] - first;
        uint32_t late = 0xffffffffu;
        for (uint32_t k = 0; k < n; ++k)
          if ((s.terms[first + k] >> 16) == prev_out)
            late = k;
        best.headers.push_back(s.tasks[2 * q]);
        best.headers.push_back(n | ((late & 0xffffu) << 16));
        for (uint32_t k = 0; k < static_cast<uint32_t>(kChainLanes); ++k)
          best.lane_terms.push_back(k < n ? s.terms[first + k] : 0u);
        prev_out = s.tasks[2 * q] & 0xffffu;
      }
      best.headers.resize(2 * kChainTasks, 0u);
      return best;
    }

    void build_cooperative_schedule(MechanismTables& t)
    {
      const int32_t n = static_cast<int32_t>(t.n_species);
      // in-place pattern: union of L (strict) and U, row by row
      std::vector<std::vector<int32_t>> rows(n);
      for (int32_t i = 0; i < n; ++i)
      {
        for (int32_t p = t.lower_row_ptr[i]; p < t.lower_row_ptr[i + 1]; ++p)
          if (t.lower_col[p] < i)
            rows[i].push_back(t.lower_col[p]);
        for (int32_t p = t.upper_row_ptr[i]; p < t.upper_row_ptr[i + 1]; ++p)
          rows[i].push_back(t.upper_col[p]);
      }
      to_csr(rows, t.alu_row_ptr, t.alu_col);
      auto alu = [&](int32_t r, int32_t c) { return position(t.alu_row_ptr, t.alu_col, r, c); };
      t.alu_diag.clear();
      for (int32_t i = 0; i < n; ++i)
        t.alu_diag.push_back(alu(i, i));
      t.jac_to_alu.clear();
      for (int32_t i = 0; i < n; ++i)
        for (int32_t p = t.jac_row_ptr[i]; p < t.jac_row_ptr[i + 1]; ++p)
          t.jac_to_alu.push_back(alu(i, t.jac_col[p]));
      t.alu_lower.assign(t.lower_col.size(), -1);
      t.alu_upper.assign(t.upper_col.size(), -1);
      for (int32_t i = 0; i < n; ++i)
      {
        for (int32_t p = t.lower_row_ptr[i]; p < t.lower_row_ptr[i + 1]; ++p)
          if (t.lower_col[p] < i)
            t.alu_lower[p] = alu(i, t.lower_col[p]);
        for (int32_t p = t.upper_row_ptr[i]; p < t.upper_row_ptr[i + 1]; ++p)
          t.alu_upper[p] = alu(i, t.upper_col[p]);
      }
      // level schedules (16-bit positions; 0xffff marks "no divisor")
      t.coop_levels_ok = t.alu_col.size() < kUnitDivisor && n < 0xffff && t.n_reactions < 0xffff;
      t.lv_lu = LevelSchedule{};
      t.lv_lu_blocked.clear();
      t.lv_lu_capped.clear();
      t.lv_fwd = LevelSchedule{};
      t.fwd_sweep = ColumnSweep{};
      t.bwd_chain = TaskChain{};
      t.lv_bwd = LevelSchedule{};
      if (!t.coop_levels_ok)
        return;
      // LU: one task per Doolittle element that has products to subtract or a divisor;
      // the level of an ALU position is one more than the highest level it reads.  One
      // schedule per blocking.
      t.lv_lu_blocked.clear();
      for (int b = 0; b < kNumLuBlockings; ++b)
      {
        const size_t min_len = kLuBlockings[b] > 0 ? static_cast<size_t>(kLuBlockings[b]) : ~size_t(0);
        std::vector<int32_t> level(t.alu_col.size(), 0), ready;
        std::vector<uint32_t> terms;
        std::vector<LevelTask> tasks;
        const size_t n_el = t.lu_out.size();
        for (size_t e = 0; e < n_el; ++e)
        {
          const bool lower = t.lu_div[e] >= 0;
          const int32_t out = lower ? t.alu_lower[t.lu_out[e]] : t.alu_upper[t.lu_out[e]];
          const int32_t tb = t.lu_term_offsets[e], te = t.lu_term_offsets[e + 1];
          if (tb == te && !lower)
            continue;  // U element equal to its Jacobian entry
          const uint32_t diag = lower ? static_cast<uint32_t>(t.alu_upper[t.lu_div[e]]) : kNoDivisor;
          terms.clear();
          ready.clear();
          for (int32_t q = tb; q < te; ++q)
          {
            const int32_t l = t.alu_lower[t.lu_term_l[q]], u = t.alu_upper[t.lu_term_u[q]];
            terms.push_back(static_cast<uint32_t>(l) | (static_cast<uint32_t>(u) << 16));
            ready.push_back(std::max(level[l], level[u]));
          }
          level[out] = add_segments(tasks, static_cast<uint32_t>(out), diag, terms, ready, lower ? level[diag] : 0,
                                    min_len) + 1;
        }
        // record slots past a task's last term name the zero kept one past the factor
        // (both operands), so the LU's records subtract 0 * 0 there
        const uint32_t pad = static_cast<uint32_t>(t.alu_col.size()) * 0x10001u;
        const TaskOrder order = task_order(tasks);
        for (int c = 0; c < kNumLuCaps; ++c)
        {
          if (order.widest <= static_cast<size_t>(kLuCaps[c]))
          {
            t.lv_lu_capped.emplace_back();
            continue;
          }
          const std::vector<int32_t> level_capped = cap_levels(tasks, order, static_cast<size_t>(kLuCaps[c]));
          t.lv_lu_capped.push_back(pack_levels(tasks, &level_capped, pad));
        }
        t.lv_lu_blocked.push_back(pack_levels(tasks, nullptr, pad));
        if (kLuBlockings[b] == kLuBlockingDefault)
          t.lv_lu = t.lv_lu_blocked.back();
      }
      // forward substitution: row i subtracts L[i][j] x[j], j ascending (unit diagonal)
      {
        std::vector<int32_t> level(n, 0), ready;
        std::vector<uint32_t> terms;
        std::vector<LevelTask> tasks;
        for (int32_t i = 0; i < n; ++i)
        {
          const int32_t b = t.fwd_offsets[i], e = t.fwd_offsets[i + 1];
          if (b == e)
            continue;
          terms.clear();
          ready.clear();
          for (int32_t q = b; q < e; ++q)
          {
            const int32_t j = t.fwd_col[q];
            terms.push_back(static_cast<uint32_t>(t.alu_lower[t.fwd_mat[q]]) | (static_cast<uint32_t>(j) << 16));
            ready.push_back(level[j]);
          }
          level[i] = add_segments(tasks, static_cast<uint32_t>(i), kUnitDivisor, terms, ready, 0, kSolveSegmentMin) + 1;
        }
        t.lv_fwd = pack_levels(tasks);
        t.fwd_sweep = column_sweep(t.lv_fwd);
      }
      // back substitution: row i subtracts U[i][j] x[j], j ascending, then divides by U[i][i]
      {
        std::vector<int32_t> level(n, 0), ready;
        std::vector<uint32_t> terms;
        std::vector<LevelTask> tasks;
        for (int32_t i = n - 1; i >= 0; --i)
        {
          terms.clear();
          ready.clear();
          for (int32_t q = t.bwd_offsets[i]; q < t.bwd_offsets[i + 1]; ++q)
          {
            const int32_t j = t.bwd_col[q];
            terms.push_back(static_cast<uint32_t>(t.alu_upper[t.bwd_mat[q]]) | (static_cast<uint32_t>(j) << 16));
            ready.push_back(level[j]);
          }
          level[i] = add_segments(tasks, static_cast<uint32_t>(i), static_cast<uint32_t>(t.alu_diag[i]), terms, ready, 0,
                                kSolveSegmentMin) + 1;
        }
        t.lv_bwd = pack_levels(tasks);
        t.bwd_chain = task_chain(t.lv_bwd);
      }
      // forcing gather, in the scatter order of the reference (reactions ascending;
      // within a reaction its reactant slots, then its product slots)
      std::vector<std::vector<std::pair<int32_t, double>>> spec(n);
      for (int32_t r = 0; r < (int32_t)t.n_reactions; ++r)
      {
        for (int32_t i = t.reactant_offsets[r]; i < t.reactant_offsets[r + 1]; ++i)
          spec[t.reactant_ids[i]].emplace_back(r, -1.0);
        for (int32_t i = t.product_offsets[r]; i < t.product_offsets[r + 1]; ++i)
          spec[t.product_ids[i]].emplace_back(r, t.yields[i]);
      }
      t.spec_offsets.assign(1, 0);
      t.spec_rxn.clear();
      t.spec_coef.clear();
      for (int32_t s = 0; s < n; ++s)
      {
        for (auto [r, cf] : spec[s])
        {
          t.spec_rxn.push_back(r);
          t.spec_coef.push_back(cf);
        }
        t.spec_offsets.push_back(static_cast<int32_t>(t.spec_rxn.size()));
      }
      // Jacobian gather in term order (deps +1, independent +1, products -yield)
      std::vector<std::vector<std::pair<int32_t, double>>> jel(t.jac_col.size());
      for (int32_t term = 0; term < (int32_t)t.term_process.size(); ++term)
      {
        const int32_t p = t.term_process[term];
        int32_t f = t.term_flat_offsets[term];
        for (int32_t i = t.term_dep_offsets[term]; i < t.term_dep_offsets[term + 1]; ++i)
          jel[t.term_flat[f++]].emplace_back(term, 1.0);
        jel[t.term_flat[f++]].emplace_back(term, 1.0);
        for (int32_t i = t.product_offsets[p]; i < t.product_offsets[p + 1]; ++i)
          jel[t.term_flat[f++]].emplace_back(term, -t.yields[i]);
      }
      t.jel_offsets.assign(1, 0);
      t.jel_term.clear();
      t.jel_coef.clear();
      for (auto& v : jel)
      {
        for (auto [tt, cf] : v)
        {
          t.jel_term.push_back(tt);
          t.jel_coef.push_back(cf);
        }
        t.jel_offsets.push_back(static_cast<int32_t>(t.jel_term.size()));
      }
      // eight padding entries let the gathers read in batches of eight past a list's end
      t.spec_rxn.insert(t.spec_rxn.end(), 8, 0);
      t.spec_coef.insert(t.spec_coef.end(), 8, 0.0);
      t.jel_term.insert(t.jel_term.end(), 8, 0);
      t.jel_coef.insert(t.jel_coef.end(), 8, 0.0);
      // packed operand lists: up to three 16-bit ids and their count (0xffff: use the lists)
      auto pack = [](const int32_t* ids, int32_t count, int32_t max_ids) {
        if (count > max_ids)
          return static_cast<uint64_t>(0xffff) << 48;
        uint64_t v = static_cast<uint64_t>(count) << 48;
        for (int32_t i = 0; i < count; ++i)
          v |= static_cast<uint64_t>(static_cast<uint16_t>(ids[i])) << (16 * i);
        return v;
      };
      t.rxn_pack.clear();
      for (int32_t r = 0; r < (int32_t)t.n_reactions; ++r)
        t.rxn_pack.push_back(pack(t.reactant_ids.data() + t.reactant_offsets[r],
                                  t.reactant_offsets[r + 1] - t.reactant_offsets[r], 3));
      t.term_pack.clear();
      for (size_t term = 0; term < t.term_process.size(); ++term)
      {
        const int32_t b = t.term_dep_offsets[term], e = t.term_dep_offsets[term + 1];
        uint64_t v = pack(t.term_dep_ids.data() + b, e - b, 2);
        v = (v & 0xffff000000000000ull) | ((v & 0xffffffffull) << 16) | static_cast<uint16_t>(t.term_process[term]);
        t.term_pack.push_back(v);
      }
      // the operands of each Jacobian gather entry, so that a gather reads one word per entry
      t.jel_pack.clear();
      for (size_t q = 0; q + 8 < t.jel_term.size(); ++q)
        t.jel_pack.push_back(t.term_pack[t.jel_term[q]]);
      t.jel_pack.insert(t.jel_pack.end(), 8, 0);
    }

    // Tables of the discrete adjoint.  The transposed substitutions have no counterpart in
    // the reference; their sums run in ascending column order.
    void build_adjoint_tables(MechanismTables& t)
    {
      const int32_t n = static_cast<int32_t>(t.n_species);
      const int32_t R = static_cast<int32_t>(t.n_reactions);
      const int32_t nt = static_cast<int32_t>(t.term_process.size());
      t.lv_ut = LevelSchedule{};
      t.lv_lt = LevelSchedule{};
      if (t.coop_levels_ok)
      {
        // columns of U and of the strict lower L, ascending rows within a column
        std::vector<std::vector<std::pair<int32_t, int32_t>>> ucol(n), lcol(n);  // (row j, ALU position)
        for (int32_t j = 0; j < n; ++j)
        {
          for (int32_t p = t.upper_row_ptr[j]; p < t.upper_row_ptr[j + 1]; ++p)
            if (t.upper_col[p] > j)
              ucol[t.upper_col[p]].emplace_back(j, t.alu_upper[p]);
          for (int32_t p = t.lower_row_ptr[j]; p < t.lower_row_ptr[j + 1]; ++p)
            if (t.lower_col[p] < j)
              lcol[t.lower_col[p]].emplace_back(j, t.alu_lower[p]);
        }
        std::vector<int32_t> level(n, 0), ready;
        std::vector<uint32_t> terms;
        std::vector<LevelTask> tasks;
        for (int32_t i = 0; i < n; ++i)
        {
          terms.clear();
          ready.clear();
          for (auto [j, pos] : ucol[i])
          {
            terms.push_back(static_cast<uint32_t>(pos) | (static_cast<uint32_t>(j) << 16));
            ready.push_back(level[j]);
          }
          level[i] = add_segments(tasks, static_cast<uint32_t>(i), static_cast<uint32_t>(t.alu_diag[i]), terms, ready, 0,
                                  kSolveSegmentMin) + 1;
        }
        t.lv_ut = pack_levels(tasks);
        std::fill(level.begin(), level.end(), 0);
        tasks.clear();
        for (int32_t i = n - 1; i >= 0; --i)
        {
          if (lcol[i].empty())
            continue;
          terms.clear();
          ready.clear();
          for (auto [j, pos] : lcol[i])
          {
            terms.push_back(static_cast<uint32_t>(pos) | (static_cast<uint32_t>(j) << 16));
            ready.push_back(level[j]);
          }
          level[i] = add_segments(tasks, static_cast<uint32_t>(i), kNoDivisor, terms, ready, 0, kSolveSegmentMin) + 1;
        }
        t.lv_lt = pack_levels(tasks);
      }
      std::vector<std::vector<int32_t>> by_rxn(R), by_indep(n);
      std::vector<std::vector<std::pair<int32_t, int32_t>>> by_dep(n);
      for (int32_t q = 0; q < nt; ++q)
      {
        by_rxn[t.term_process[q]].push_back(q);
        by_indep[t.term_indep[q]].push_back(q);
        for (int32_t a = t.term_dep_offsets[q]; a < t.term_dep_offsets[q + 1]; ++a)
          by_dep[t.term_dep_ids[a]].emplace_back(q, a - t.term_dep_offsets[q]);
      }
      auto flatten = [](const std::vector<std::vector<int32_t>>& v, std::vector<int32_t>& off, std::vector<int32_t>& out) {
        off.assign(1, 0);
        out.clear();
        for (const auto& l : v)
        {
          out.insert(out.end(), l.begin(), l.end());
          off.push_back(static_cast<int32_t>(out.size()));
        }
      };
      flatten(by_rxn, t.rxn_term_offsets, t.rxn_terms);
      flatten(by_indep, t.indep_term_offsets, t.indep_terms);
      t.hes_offsets.assign(1, 0);
      t.hes_term.clear();
      t.hes_slot.clear();
      for (const auto& l : by_dep)
      {
        for (auto [q, a] : l)
        {
          t.hes_term.push_back(q);
          t.hes_slot.push_back(a);
        }
        t.hes_offsets.push_back(static_cast<int32_t>(t.hes_term.size()));
      }
    }
  }  // namespace

  MechanismTables compile_mechanism(const micm_hip_mechanism_desc& d)
  {
    if (d.n_species <= 0)
      throw std::invalid_argument("Provided chemical system contains no species.");
    if (d.n_reactions < 0 || (d.n_reactions > 0 && (!d.reactant_offsets || !d.product_offsets || !d.rate_types ||
                                                    !d.rate_params)))
      throw std::invalid_argument("incomplete mechanism description");
    if (d.n_species > (1 << 20) || d.n_reactions > (1 << 24))
      throw std::invalid_argument("mechanism too large");
    const int32_t n = static_cast<int32_t>(d.n_species);
    const int64_t R = d.n_reactions;
    MechanismTables t;
    t.n_species = n;
    t.n_reactions = R;

    // ---- validate species references
    for (int64_t r = 0; r < R; ++r)
    {
      if (d.rate_types[r] < 0 || d.rate_types[r] >= MICM_HIP_RATE_N_TYPES)
        throw std::invalid_argument("unknown rate-constant type");
      for (int64_t i = d.reactant_offsets[r]; i < d.reactant_offsets[r + 1]; ++i)
        if (d.reactants[i] >= n || -d.reactants[i] - 1 >= d.n_parameterized)
          throw std::domain_error("reactant does not exist");
      for (int64_t i = d.product_offsets[r]; i < d.product_offsets[r + 1]; ++i)
        if (d.products[i] >= n || -d.products[i] - 1 >= d.n_parameterized)
          throw std::domain_error("product does not exist");
    }
    if (!d.ignore_unused_species)
    {
      std::vector<char> used(n, 0);
      for (int64_t r = 0; r < R; ++r)
      {
        for (int64_t i = d.reactant_offsets[r]; i < d.reactant_offsets[r + 1]; ++i)
          if (d.reactants[i] >= 0)
            used[d.reactants[i]] = 1;
        for (int64_t i = d.product_offsets[r]; i < d.product_offsets[r + 1]; ++i)
          if (d.products[i] >= 0)
            used[d.products[i]] = 1;
      }
      for (int32_t s = 0; s < n; ++s)
        if (!used[s])
          throw std::domain_error("Unused species in chemical system");
    }

    // ---- reactions stable-sorted by type
    t.sorted_to_declared.resize(R);
    for (int64_t r = 0; r < R; ++r)
      t.sorted_to_declared[r] = r;
    std::stable_sort(t.sorted_to_declared.begin(), t.sorted_to_declared.end(),
                     [&](int64_t a, int64_t b) { return d.rate_types[a] < d.rate_types[b]; });

    // ---- Jacobian non-zeros in declared indices, then the species order
    std::vector<std::pair<int32_t, int32_t>> nz;
    for (int64_t r = 0; r < R; ++r)
    {
      for (int64_t i = d.reactant_offsets[r]; i < d.reactant_offsets[r + 1]; ++i)
      {
        const int64_t ind = d.reactants[i];
        if (ind < 0)
          continue;
        for (int64_t j = d.reactant_offsets[r]; j < d.reactant_offsets[r + 1]; ++j)
          if (d.reactants[j] >= 0)
            nz.emplace_back(static_cast<int32_t>(d.reactants[j]), static_cast<int32_t>(ind));
        for (int64_t j = d.product_offsets[r]; j < d.product_offsets[r + 1]; ++j)
          if (d.products[j] >= 0)
            nz.emplace_back(static_cast<int32_t>(d.products[j]), static_cast<int32_t>(ind));
      }
    }
    std::sort(nz.begin(), nz.end());
    nz.erase(std::unique(nz.begin(), nz.end()), nz.end());
    if (d.reorder_state)
      t.state_to_declared = markowitz_order(n, nz);
    else
    {
      t.state_to_declared.resize(n);
      for (int32_t i = 0; i < n; ++i)
        t.state_to_declared[i] = i;
    }
    std::vector<int32_t> to_state(n);
    for (int32_t i = 0; i < n; ++i)
      to_state[t.state_to_declared[i]] = i;
    t.absolute_tolerances.assign(n, 1e-3);
    if (d.absolute_tolerances)
      for (int32_t i = 0; i < n; ++i)
        t.absolute_tolerances[i] = d.absolute_tolerances[t.state_to_declared[i]];

    std::vector<std::vector<int32_t>> arow(n);
    for (auto [r, c] : nz)
      arow[to_state[r]].push_back(to_state[c]);
    for (int32_t i = 0; i < n; ++i)
    {
      arow[i].push_back(i);
      std::sort(arow[i].begin(), arow[i].end());
      arow[i].erase(std::unique(arow[i].begin(), arow[i].end()), arow[i].end());
    }
    to_csr(arow, t.jac_row_ptr, t.jac_col);

    // ---- rate-constant store and ProcessSet tables (sorted reactions, state indices)
    t.multiplier_offsets.push_back(0);
    t.reactant_offsets.push_back(0);
    t.product_offsets.push_back(0);
    int32_t custom = 0;
    for (int64_t k = 0; k < R; ++k)
    {
      const int64_t r = t.sorted_to_declared[k];
      const int32_t type = d.rate_types[r];
      t.rate_type.push_back(type);
      std::array<double, MICM_HIP_RATE_PARAM_STRIDE> p{};
      for (int j = 0; j < MICM_HIP_RATE_PARAM_STRIDE; ++j)
        p[j] = d.rate_params[r * MICM_HIP_RATE_PARAM_STRIDE + j];
      if (type == MICM_HIP_RATE_BRANCHED)
      {
        const double k0 = 2.0e-22 * kAvogadro * 1.0e-6 * std::exp(static_cast<double>(static_cast<int64_t>(p[4])));
        const double air_ref = 2.45e19 / kAvogadro * 1.0e6;
        const double a = k0 * air_ref;
        const double b = 0.43 * std::pow(293.0 / 298.0, -8.0);
        const double A_val = a / (1.0 + a / b) * std::pow(0.41, 1.0 / (1.0 + std::pow(std::log10(a / b), 2.0)));
        p[5] = k0;
        p[6] = A_val * (1.0 - p[3]) / p[3];
      }
      if (type == MICM_HIP_RATE_SURFACE)
        p[3] = 8.0 * kGasConstant / (kPi * p[1]);
      if (type == MICM_HIP_RATE_TAYLOR && (p[5] < 0 || p[5] > 16))
        throw std::invalid_argument("Taylor series rate constants take at most 16 coefficients");
      t.rate_params.insert(t.rate_params.end(), p.begin(), p.end());
      if (type == MICM_HIP_RATE_USER_DEFINED)
      {
        t.custom_index.push_back(custom);
        t.custom_owner.push_back(k);
        custom += 1;
      }
      else if (type == MICM_HIP_RATE_SURFACE)
      {
        t.custom_index.push_back(custom);
        t.custom_owner.push_back(k);
        t.custom_owner.push_back(k);
        custom += 2;
      }
      else
        t.custom_index.push_back(-1);
      for (int64_t i = d.reactant_offsets[r]; i < d.reactant_offsets[r + 1]; ++i)
      {
        const int64_t id = d.reactants[i];
        if (id >= 0)
          t.reactant_ids.push_back(to_state[id]);
        else
          for (int j = 0; j < 4; ++j)
            t.multipliers.push_back(d.parameterized[4 * (-id - 1) + j]);
      }
      t.reactant_offsets.push_back(static_cast<int32_t>(t.reactant_ids.size()));
      t.multiplier_offsets.push_back(static_cast<int32_t>(t.multipliers.size() / 4));
      for (int64_t i = d.product_offsets[r]; i < d.product_offsets[r + 1]; ++i)
        if (d.products[i] >= 0)
        {
          t.product_ids.push_back(to_state[d.products[i]]);
          t.yields.push_back(d.yields[i]);
        }
      t.product_offsets.push_back(static_cast<int32_t>(t.product_ids.size()));
    }
    t.n_custom = custom;

    // ---- Jacobian terms: ordered by independent species, then by process (stable)
    std::vector<std::pair<int32_t, int32_t>> columns;  // (independent, process)
    for (int32_t k = 0; k < (int32_t)R; ++k)
      for (int32_t i = t.reactant_offsets[k]; i < t.reactant_offsets[k + 1]; ++i)
        columns.emplace_back(t.reactant_ids[i], k);
    std::stable_sort(columns.begin(), columns.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    t.term_dep_offsets.push_back(0);
    t.term_flat_offsets.push_back(0);
    for (auto [ind, k] : columns)
    {
      t.term_process.push_back(k);
      t.term_indep.push_back(ind);
      bool skipped = false;
      for (int32_t i = t.reactant_offsets[k]; i < t.reactant_offsets[k + 1]; ++i)
      {
        const int32_t s = t.reactant_ids[i];
        if (s == ind && !skipped)
        {
          skipped = true;
          continue;
        }
        t.term_dep_ids.push_back(s);
        t.term_flat.push_back(position(t.jac_row_ptr, t.jac_col, s, ind));
      }
      t.term_dep_offsets.push_back(static_cast<int32_t>(t.term_dep_ids.size()));
      t.term_flat.push_back(position(t.jac_row_ptr, t.jac_col, ind, ind));
      for (int32_t i = t.product_offsets[k]; i < t.product_offsets[k + 1]; ++i)
        t.term_flat.push_back(position(t.jac_row_ptr, t.jac_col, t.product_ids[i], ind));
      t.term_flat_offsets.push_back(static_cast<int32_t>(t.term_flat.size()));
    }

    // ---- Doolittle fill-in, one row of U and one column of L at a time
    std::vector<std::vector<int32_t>> acol(n);
    for (int32_t i = 0; i < n; ++i)
      for (int32_t c : arow[i])
        acol[c].push_back(i);
    std::vector<std::vector<int32_t>> urow(n), ucol(n), lrow(n), lcol(n);  // strictly-lower L in lrow/lcol
    std::vector<char> mark(n, 0);
    std::vector<int32_t> touched;
    for (int32_t i = 0; i < n; ++i)
    {
      touched.clear();
      auto add = [&](int32_t k)
      {
        if (!mark[k])
        {
          mark[k] = 1;
          touched.push_back(k);
        }
      };
      add(i);
      for (int32_t k : arow[i])
        if (k >= i)
          add(k);
      for (int32_t j : lrow[i])
        for (int32_t k : urow[j])
          if (k >= i)
            add(k);
      std::sort(touched.begin(), touched.end());
      for (int32_t k : touched)
      {
        urow[i].push_back(k);
        ucol[k].push_back(i);
        mark[k] = 0;
      }
      touched.clear();
      for (int32_t k : acol[i])
        if (k > i)
          add(k);
      for (int32_t j : ucol[i])
        if (j < i)
          for (int32_t k : lcol[j])
            if (k > i)
              add(k);
      std::sort(touched.begin(), touched.end());
      for (int32_t k : touched)
      {
        lcol[i].push_back(k);
        lrow[k].push_back(i);
        mark[k] = 0;
      }
    }
    std::vector<std::vector<int32_t>> lrow_full(n);
    for (int32_t i = 0; i < n; ++i)
    {
      lrow_full[i] = lrow[i];
      lrow_full[i].push_back(i);
    }
    to_csr(lrow_full, t.lower_row_ptr, t.lower_col);
    to_csr(urow, t.upper_row_ptr, t.upper_col);
    for (int32_t i = 0; i < n; ++i)
    {
      t.jac_diag.push_back(position(t.jac_row_ptr, t.jac_col, i, i));
      t.lower_diag.push_back(position(t.lower_row_ptr, t.lower_col, i, i));
      t.upper_diag.push_back(position(t.upper_row_ptr, t.upper_col, i, i));
    }

    // ---- Doolittle schedule
    t.lu_term_offsets.push_back(0);
    for (int32_t i = 0; i < n; ++i)
    {
      for (int32_t k : urow[i])
      {
        for (int32_t j : lrow[i])
          if (has(urow[j], k))
          {
            t.lu_term_l.push_back(position(t.lower_row_ptr, t.lower_col, i, j));
            t.lu_term_u.push_back(position(t.upper_row_ptr, t.upper_col, j, k));
          }
        t.lu_out.push_back(position(t.upper_row_ptr, t.upper_col, i, k));
        t.lu_a.push_back(position(t.jac_row_ptr, t.jac_col, i, k));
        t.lu_div.push_back(-1);
        t.lu_term_offsets.push_back(static_cast<int32_t>(t.lu_term_l.size()));
      }
      for (int32_t k : lcol[i])
      {
        for (int32_t j : lrow[k])
        {
          if (j >= i)
            break;
          if (has(urow[j], i))
          {
            t.lu_term_l.push_back(position(t.lower_row_ptr, t.lower_col, k, j));
            t.lu_term_u.push_back(position(t.upper_row_ptr, t.upper_col, j, i));
          }
        }
        t.lu_out.push_back(position(t.lower_row_ptr, t.lower_col, k, i));
        t.lu_a.push_back(position(t.jac_row_ptr, t.jac_col, k, i));
        t.lu_div.push_back(t.upper_diag[i]);
        t.lu_term_offsets.push_back(static_cast<int32_t>(t.lu_term_l.size()));
      }
    }

    // ---- substitution tables
    t.fwd_offsets.push_back(0);
    t.bwd_offsets.push_back(0);
    for (int32_t i = 0; i < n; ++i)
    {
      for (int32_t j : lrow[i])
      {
        t.fwd_mat.push_back(position(t.lower_row_ptr, t.lower_col, i, j));
        t.fwd_col.push_back(j);
      }
      t.fwd_offsets.push_back(static_cast<int32_t>(t.fwd_mat.size()));
      for (int32_t j : urow[i])
        if (j > i)
        {
          t.bwd_mat.push_back(position(t.upper_row_ptr, t.upper_col, i, j));
          t.bwd_col.push_back(j);
        }
      t.bwd_offsets.push_back(static_cast<int32_t>(t.bwd_mat.size()));
    }
    build_cooperative_schedule(t);
    build_adjoint_tables(t);
    for (size_t i = 0; i < t.lu_term_l.size(); ++i)
    {
      t.lu_term_pairs.push_back(t.lu_term_l[i]);
      t.lu_term_pairs.push_back(t.lu_term_u[i]);
    }
    for (size_t i = 0; i < t.fwd_mat.size(); ++i)
    {
      t.fwd_pairs.push_back(t.fwd_mat[i]);
      t.fwd_pairs.push_back(t.fwd_col[i]);
    }
    for (size_t i = 0; i < t.bwd_mat.size(); ++i)
    {
      t.bwd_pairs.push_back(t.bwd_mat[i]);
      t.bwd_pairs.push_back(t.bwd_col[i]);
    }
    return t;
  }

  bool rosenbrock_preset(int32_t preset, micm_hip_rosenbrock_parameters& p)
  {
    p = micm_hip_rosenbrock_parameters{};
    p.max_number_of_steps = 1000;
    p.round_off = std::numeric_limits<double>::epsilon();
    p.factor_min = 0.2;
    p.factor_max = 6;
    p.rejection_factor_decrease = 0.1;
    p.safety_factor = 0.9;
    // Tableaus: Hairer & Wanner, Solving ODEs II (2-, 4- and 6-stage), Sandu et al.,
    // Atmos. Environ. 31 (1997) 3459 (3-stage); values as rosenbrock_solver_parameters.hpp:152-433.
    switch (preset)
    {
      case MICM_HIP_ROS_TWO_STAGE:
      {
        const double g = 1.0 + 1.0 / 1.41421356237309504880;
        p.stages = 2;
        p.a[0] = 1.0 / g;
        p.c[0] = -2.0 / g;
        for (int i = 0; i < 6; ++i)
          p.new_function_evaluation[i] = 1;
        p.m[0] = 3.0 / (2.0 * g);
        p.m[1] = 1.0 / (2.0 * g);
        p.e[0] = 1.0 / (2.0 * g);
        p.e[1] = 1.0 / (2.0 * g);
        p.estimator_of_local_order = 2.0;
        p.alpha[1] = 1.0;
        p.gamma[0] = g;
        p.gamma[1] = -g;
        return true;
      }
      case MICM_HIP_ROS_THREE_STAGE:
        p.stages = 3;
        p.a[0] = 1;
        p.a[1] = 1;
        p.c[0] = -0.10156171083877702091975600115545e+01;
        p.c[1] = 0.40759956452537699824805835358067e+01;
        p.c[2] = 0.92076794298330791242156818474003e+01;
        p.new_function_evaluation[0] = p.new_function_evaluation[1] = 1;
        p.m[0] = 0.1e+01;
        p.m[1] = 0.61697947043828245592553615689730e+01;
        p.m[2] = -0.42772256543218573326238373806514;
        p.e[0] = 0.5;
        p.e[1] = -0.29079558716805469821718236208017e+01;
        p.e[2] = 0.22354069897811569627360909276199;
        p.estimator_of_local_order = 3;
        p.alpha[1] = p.alpha[2] = 0.43586652150845899941601945119356;
        p.gamma[0] = 0.43586652150845899941601945119356;
        p.gamma[1] = 0.24291996454816804366592249683314;
        p.gamma[2] = 0.21851380027664058511513169485832e+01;
        return true;
      case MICM_HIP_ROS_FOUR_STAGE:
        p.stages = 4;
        p.a[0] = 0.2000000000000000E+01;
        p.a[1] = 0.1867943637803922E+01;
        p.a[2] = 0.2344449711399156;
        p.a[3] = p.a[1];
        p.a[4] = p.a[2];
        p.c[0] = -0.7137615036412310E+01;
        p.c[1] = 0.2580708087951457E+01;
        p.c[2] = 0.6515950076447975;
        p.c[3] = -0.2137148994382534E+01;
        p.c[4] = -0.3214669691237626;
        p.c[5] = -0.6949742501781779;
        p.new_function_evaluation[0] = p.new_function_evaluation[1] = p.new_function_evaluation[2] = 1;
        p.m[0] = 0.2255570073418735E+01;
        p.m[1] = 0.2870493262186792;
        p.m[2] = 0.4353179431840180;
        p.m[3] = 0.1093502252409163E+01;
        p.e[0] = -0.2815431932141155;
        p.e[1] = -0.7276199124938920E-01;
        p.e[2] = -0.1082196201495311;
        p.e[3] = -0.1093502252409163E+01;
        p.estimator_of_local_order = 4.0;
        p.alpha[1] = 0.1145640000000000E+01;
        p.alpha[2] = p.alpha[3] = 0.6552168638155900;
        p.gamma[0] = 0.5728200000000000;
        p.gamma[1] = -0.1769193891319233E+01;
        p.gamma[2] = 0.7592633437920482;
        p.gamma[3] = -0.1049021087100450;
        return true;
      case MICM_HIP_ROS_FOUR_STAGE_DA:
        p.stages = 4;
        p.a[1] = 2.0;
        p.a[3] = 2.0;
        p.a[5] = 1.0;
        p.c[0] = 4.0;
        p.c[1] = 1.0;
        p.c[2] = -1.0;
        p.c[3] = 1.0;
        p.c[4] = -1.0;
        p.c[5] = -(8.0 / 3.0);
        p.new_function_evaluation[0] = p.new_function_evaluation[2] = p.new_function_evaluation[3] = 1;
        p.m[0] = 2.0;
        p.m[2] = 1.0;
        p.m[3] = 1.0;
        p.e[3] = 1.0;
        p.estimator_of_local_order = 3.0;
        p.alpha[2] = p.alpha[3] = 1.0;
        p.gamma[0] = 0.5;
        p.gamma[1] = 1.5;
        return true;
      case MICM_HIP_ROS_SIX_STAGE_DA:
      {
        p.stages = 6;
        const double al[6] = { 0.000, 0.386, 0.210, 0.630, 1.000, 1.000 };
        const double ga[6] = { 0.2500000000000000, -0.1043000000000000, 0.1035000000000000, -0.3620000000000023E-01, 0.0, 0.0 };
        const double a[15] = { 0.1544000000000000E+01, 0.9466785280815826, 0.2557011698983284, 0.3314825187068521E+01,
                               0.2896124015972201E+01, 0.9986419139977817, 0.1221224509226641E+01, 0.6019134481288629E+01,
                               0.1253708332932087E+02, -0.6878860361058950, 0.1221224509226641E+01, 0.6019134481288629E+01,
                               0.1253708332932087E+02, -0.6878860361058950, 1.0 };
        const double c[15] = { -0.5668800000000000E+01, -0.2430093356833875E+01, -0.2063599157091915,
                               -0.1073529058151375,     -0.9594562251023355E+01, -0.2047028614809616E+02,
                               0.7496443313967647E+01,  -0.1024680431464352E+02, -0.3399990352819905E+02,
                               0.1170890893206160E+02,  0.8083246795921522E+01,  -0.7981132988064893E+01,
                               -0.3152159432874371E+02, 0.1631930543123136E+02,  -0.6058818238834054E+01 };
        for (int i = 0; i < 6; ++i)
        {
          p.alpha[i] = al[i];
          p.gamma[i] = ga[i];
          p.new_function_evaluation[i] = 1;
        }
        for (int i = 0; i < 15; ++i)
        {
          p.a[i] = a[i];
          p.c[i] = c[i];
        }
        p.m[0] = a[6];
        p.m[1] = a[7];
        p.m[2] = a[8];
        p.m[3] = a[9];
        p.m[4] = 1.0;
        p.m[5] = 1.0;
        p.e[5] = 1.0;
        p.estimator_of_local_order = 4.0;
        return true;
      }
    }
    return false;
  }
}  // namespace micm_hip
