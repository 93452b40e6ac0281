// mechanism.hpp -- host-side compilation of a mechanism into the index tables
// the HIP kernels consume.  Replaces the symbolic half of the reference's
// SolverBuilder::Build (include/micm/solver/solver_builder.inl:227-548).
#pragma once

#include "../../include/micm_hip.h"

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace micm_hip
{
  // Divisor field of a level task: an ALU position, or one of these markers.
  constexpr uint32_t kNoDivisor = 0xffff;    // no division
  constexpr uint32_t kUnitDivisor = 0xfffe;  // forward substitution: divide by the row's L[i][i]

  // Blocking of the LU level schedule: the shortest segment (products) an element's
  // ordered sum is split into, except its last; 0 keeps each sum whole.  Longer segments
  // mean fewer tasks per level, shorter ones a sum that starts as soon as its first
  // operands are final.  Results do not depend on it.
  constexpr int kLuBlockings[] = { 1, 2, 4, 8, 0 };
  constexpr int kNumLuBlockings = 5;
  constexpr int kLuBlockingDefault = 4;

  // Widths of the LU schedules with capped levels: a workgroup of W threads runs the
  // schedule whose levels hold at most W tasks, so that every thread has at most one task
  // per level (the one whose record it read ahead); wider workgroups run the uncapped
  // schedule.  Index c of lv_lu_capped.
  constexpr int kLuCaps[] = { 64, 128, 256, 512 };
  constexpr int kNumLuCaps = 4;

  // Terms carried in a task's record (LevelSchedule::records: 16 words per task).
  constexpr int kInlineTerms = 13;
  constexpr int kRecordWords = 3 + kInlineTerms;

  // Independent tasks grouped by dependency level (see MechanismTables::lv_lu).
  struct LevelSchedule
  {
    std::vector<int32_t> offsets;  // level -> first task, n_levels + 1
    std::vector<uint32_t> tasks;   // 2 per task + sentinel: out | diag << 16, first term
    std::vector<uint32_t> terms;   // two 16-bit fields, then sixteen zero entries of padding
    // kRecordWords per task + sentinel: out | diag << 16, first term, number of terms, then
    // the task's first kInlineTerms terms, so that one 64-byte read brings a task.  Slots
    // past a task's last term hold 0, or in the LU schedules the position one past the
    // factor in both fields: the device keeps a zero there, so that a padded slot subtracts
    // 0 * 0 and needs no test.
    std::vector<uint32_t> records;
    int32_t max_terms = 0;
    int32_t max_level_terms = 0;  // largest number of terms of one level
    int32_t n_levels() const { return offsets.empty() ? 0 : static_cast<int32_t>(offsets.size()) - 1; }
    int32_t n_tasks() const { return static_cast<int32_t>(tasks.size() / 2) - 1; }
  };

  // The last levels of the forward substitution run as one column sweep by a single
  // wavefront: lane r holds the partial sum of row rows[r]; the columns (x positions the
  // swept terms read) are taken in ascending order, and at column c every lane whose row
  // has a term there subtracts L * x[c], x[c] being another lane's finished row or a value
  // final before the sweep.  Each row's terms stay in their order, so the results are
  // bit for bit those of the level schedule; the chain of one-term levels costs a few
  // instructions per column instead of a workgroup barrier and a round trip per level.
  struct ColumnSweep
  {
    int32_t first_level = -1;  // levels from here on are swept (-1: no sweep)
    int32_t n_cols = 0;
    // kSweepLanes entries of 4 words: row | start << 16 (0xffff: no row; start = ALU
    // position of the row's first swept term, the others following it), the 64-bit mask
    // of the columns the row has terms at, and this lane's column source: the lane whose
    // row is that column | kSweepFromLane, else the x position
    std::vector<uint32_t> lanes;
  };
  constexpr int kSweepLanes = 64;
  constexpr uint32_t kSweepFromLane = 0x80000000u;

  // The first levels of the back substitution as one chain of tasks run in order by a
  // single wavefront: lane k forms product k of a task, the wavefront subtracts them in
  // order.  While task t is summed, the lanes already form the products of task t + 1,
  // except the one (`late`) that reads task t's own result, which the sum of task t + 1
  // forms in its place; so the chain costs about one multiply, the task's subtractions
  // and its division per task instead of two workgroup barriers and three LDS round trips
  // per level.  The same operations in the same order as the level schedule.
  struct TaskChain
  {
    int32_t levels = 0;  // levels [0, levels) of the schedule run as the chain (0: none)
    // kChainTasks entries of 2 words (0 past the last task): out | diag << 16, number of
    // terms | index of the late term << 16 (0xffff: none)
    std::vector<uint32_t> headers;
    std::vector<uint32_t> lane_terms;  // kChainLanes per task: term k of the task (0 past its last)
  };
  constexpr int kChainLanes = 64;
  constexpr int kChainTasks = 128;  // the headers sit in two registers of each lane

  // Rate parameters as consumed by the device (per sorted reaction, stride
  // MICM_HIP_RATE_PARAM_STRIDE).  Derived values live in spare slots:
  //   BRANCHED  slot 5 = k0, slot 6 = z   (reaction_rate_store.hpp:286-293)
  //   SURFACE   slot 3 = mean free speed factor 8R/(pi*M) (reaction_rate_store.hpp:328)
  struct MechanismTables
  {
    int64_t n_species = 0, n_reactions = 0, n_custom = 0;
    std::vector<int64_t> state_to_declared;   // species permutation
    std::vector<int64_t> sorted_to_declared;  // reaction permutation
    std::vector<double> absolute_tolerances;  // state order

    // ---- rate constants (sorted reactions)
    std::vector<int32_t> rate_type;
    std::vector<double> rate_params;          // n_reactions * stride
    std::vector<int32_t> custom_index;        // first custom parameter, -1 if none
    std::vector<int32_t> multiplier_offsets;  // n_reactions + 1
    std::vector<double> multipliers;          // 4 per parameterized reactant: c0, cT, cP, crho
    std::vector<int64_t> custom_owner;        // custom parameter -> sorted reaction

    // ---- forcing (ProcessSet, process_set.hpp:280-317)
    std::vector<int32_t> reactant_offsets, reactant_ids;
    std::vector<int32_t> product_offsets, product_ids;
    std::vector<double> yields;

    // ---- Jacobian terms (process_set.hpp:329-391, flat ids :460-492)
    std::vector<int32_t> term_process;        // one per (independent reactant occurrence, process)
    std::vector<int32_t> term_dep_offsets;    // dependent reactants of each term
    std::vector<int32_t> term_dep_ids;
    std::vector<int32_t> term_flat_offsets;   // deps..., independent, products...
    std::vector<int32_t> term_flat;

    // ---- sparsity patterns (CSR, row-major sorted)
    std::vector<int32_t> jac_row_ptr, jac_col;   // with the diagonal (jacobian.hpp:27-30)
    std::vector<int32_t> lower_row_ptr, lower_col;
    std::vector<int32_t> upper_row_ptr, upper_col;
    std::vector<int32_t> jac_diag, lower_diag, upper_diag;

    // ---- Doolittle schedule (lu_decomposition_doolittle.inl:177-299): per output element
    // in the reference's order (row i: U[i][k] for k ascending, then L[k][i] for k ascending).
    std::vector<int32_t> lu_out;          // storage index in U (div < 0) or L (div >= 0)
    std::vector<int32_t> lu_a;            // storage index in the Jacobian, -1 if structurally zero
    std::vector<int32_t> lu_div;          // U[i][i] storage index for L elements, -1 for U elements
    std::vector<int32_t> lu_term_offsets; // n_elements + 1
    std::vector<int32_t> lu_term_l;       // L[i][j] (or L[k][j]) storage index, j ascending
    std::vector<int32_t> lu_term_u;       // U[j][k] (or U[j][i]) storage index
    std::vector<int32_t> lu_term_pairs;   // interleaved (l, u) of the two arrays above

    // ---- substitution (linear_solver.inl:166-224)
    std::vector<int32_t> fwd_offsets, fwd_mat, fwd_col;   // row i: L[i][j], j ascending
    std::vector<int32_t> bwd_offsets, bwd_mat, bwd_col;   // row i (ascending storage): U[i][j], j ascending
    std::vector<int32_t> fwd_pairs, bwd_pairs;            // interleaved (matrix index, column)

    // ---- cooperative (one workgroup per cell) schedule over the in-place pattern
    // ALU = strictly-lower L united with U (CSR, row-major), factorised in place.
    std::vector<int32_t> alu_row_ptr, alu_col, alu_diag;
    std::vector<int32_t> jac_to_alu;                       // Jacobian element -> ALU position
    std::vector<int32_t> alu_lower, alu_upper;             // L / U storage index -> ALU position
    // Level schedules: every task reads only final values written at lower levels and
    // subtracts its products in the reference's order, so the tasks of one level run
    // in parallel without changing any rounding.  An element's sum may be split into
    // segments at increasing levels (its partial sum kept in its own position); only
    // the last segment divides.
    bool coop_levels_ok = false;                           // positions fit the 16-bit fields
    LevelSchedule lv_lu;   // out = ALU position, diag = U[i][i] position (L elements), terms (l | u << 16)
    // the LU schedule for each shortest-segment length of kLuBlockings (lv_lu is the one
    // for kLuBlockingDefault)
    std::vector<LevelSchedule> lv_lu_blocked;
    // lv_lu_capped[b * kNumLuCaps + c]: lv_lu_blocked[b] re-levelled so that no level
    // holds more than kLuCaps[c] tasks (empty where no level of lv_lu_blocked[b] does)
    std::vector<LevelSchedule> lv_lu_capped;
    LevelSchedule lv_fwd;  // out = row, diag = kUnitDivisor on a row's last segment, terms (ALU position | column << 16)
    ColumnSweep fwd_sweep;  // lv_fwd's last levels as a column sweep
    TaskChain bwd_chain;    // lv_bwd's first levels as a task chain
    LevelSchedule lv_bwd;  // out = row, diag = U[i][i] position, terms (ALU position | column << 16)
    // packed operands: reaction -> reactant ids, Jacobian term -> (process, dependent ids);
    // three / two 16-bit ids and the count in bits 48-63 (0xffff: more, use the lists)
    std::vector<uint64_t> rxn_pack, term_pack;
    // forcing as a per-species gather: f[s] += c * rate[r] over (r, c) in reaction order
    std::vector<int32_t> spec_offsets, spec_rxn;
    std::vector<double> spec_coef;
    // -J as a per-element gather: J[e] += c * d[t] over (t, c) in the reference's order
    std::vector<int32_t> jel_offsets, jel_term;
    std::vector<double> jel_coef;
    std::vector<uint64_t> jel_pack;                        // term_pack of each gather entry

    // ---- discrete adjoint (one workgroup per cell, over the in-place factor)
    std::vector<int32_t> term_indep;  // independent species of each Jacobian term
    // transposed substitutions: A^T x = b with A = L U is U^T z = b (rows ascending), then
    // L^T x = z (rows descending); row i of U^T is column i of U
    LevelSchedule lv_ut;  // out = row, diag = U[i][i] position, terms (ALU position of U[j][i] | j << 16), j ascending
    LevelSchedule lv_lt;  // out = row, no divisor, terms (ALU position of L[j][i] | j << 16), j ascending
    std::vector<int32_t> rxn_term_offsets, rxn_terms;      // Jacobian terms of each reaction
    std::vector<int32_t> indep_term_offsets, indep_terms;  // Jacobian terms whose independent species is q
    // second derivative: entry (term t, dependent slot a) differentiates d_t by the species
    // in slot a; grouped by that species
    std::vector<int32_t> hes_offsets, hes_term, hes_slot;

    int64_t nnz_alu() const { return (int64_t)alu_col.size(); }
    int64_t nnz_jacobian() const { return (int64_t)jac_col.size(); }
    int64_t nnz_lower() const { return (int64_t)lower_col.size(); }
    int64_t nnz_upper() const { return (int64_t)upper_col.size(); }
    int64_t lu_multiply_adds() const { return (int64_t)lu_term_l.size(); }
  };

  // Throws std::invalid_argument (bad description) or std::domain_error (species errors).
  MechanismTables compile_mechanism(const micm_hip_mechanism_desc& desc);

  // RosenbrockSolverParameters presets (rosenbrock_solver_parameters.hpp:152-433).
  bool rosenbrock_preset(int32_t preset, micm_hip_rosenbrock_parameters& out);
}  // namespace micm_hip
