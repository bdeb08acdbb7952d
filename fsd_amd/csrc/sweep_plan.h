// Static schedules of the two block-triangular sweeps of G = P^T L^-T L^-1 P
// for the single-workgroup sweep kernel (host side).
#pragma once
#include <cstdint>
#include <vector>

// Lanes of the sweep workgroup: 16 waves of 64.
constexpr int TP_LANES = 1024;
constexpr int TP_WAVE = 64;
// LDS ring of recent results (rows; a power of two) and rows per step, so
// that any three consecutive steps fit in the ring
constexpr int TP_RING = 2048;
constexpr int TP_ROWS_MAX = 640;
// src field of a lane: dependency read from the global result array (the
// rest of the word is the producer's sequence number) instead of the ring
constexpr int TP_GLOBAL = 0x40000000;
constexpr int TP_ND_SHIFT = 24;   // own field: sequence number | lanes << TP_ND_SHIFT

// One sweep, cut into steps.  A step holds rows of a single level of the
// sweep's DAG, packed into the 1024 lanes: a row with nd dependencies takes
// nd consecutive lanes of one wave (one dependency block each, the row's
// first lane owns the row), a row without dependencies takes one lane.
// Rows are numbered in step order (the sequence number k); every per-row
// value (input, result) lives at index k.  Lane g = step * 1024 + threadIdx.x.
//   info[2g]     src   -1: no block; r < TP_GLOBAL: the dependency's result
//                      is in ring slot r; TP_GLOBAL | k: it is result k in
//                      global memory
//   info[2g + 1] own   owner: k | nd << TP_ND_SHIFT (nd = 0 for a row
//                      without block), else -1
//   bsrc[2g], bsrc[2g + 1]: factor block index and the particle whose
//                      diagonal block scales it (-1 when no block)
// A row's result enters the LDS ring at slot (k mod TP_RING) during its
// step, and global memory during the next step (its store is issued just
// before the next step's operands are fetched); every step ends with a
// workgroup barrier.  The operands of step S are fetched during step S - 1,
// after the barrier that ends step S - 2, which orders the global stores of
// results of steps up to S - 3.  So a dependency of step S comes from the
// ring when its producer is in steps S - 2 or S - 1 (three steps of at most
// TP_ROWS_MAX rows fit in the ring) or is otherwise still there, and from
// global memory when it is older.
// build_task_plan places the rows (seq, lane, step of every row; the end of
// every step); task_plan_lanes derives the lane arrays from the placement
// (the library does that on the GPU, tp_lanes_kernel, from the same rules).
struct TaskPlan {
    int nsteps = 0;
    std::vector<int> seq;       // permuted particle -> sequence number
    std::vector<int> lane;      // permuted particle -> its first lane (step * TP_LANES + lane in the step)
    std::vector<int> step;      // permuted particle -> its step
    std::vector<int> send;      // step -> one past the last sequence number of the step
    std::vector<int> info;      // 2 per lane (task_plan_lanes)
    std::vector<int> bsrc;      // 2 per lane (task_plan_lanes)
    std::vector<int> kptr, order, lptr;   // scratch of the placement
};

// lev[P] in [0, nlev): level of permuted particle P in the sweep's DAG;
// rowptr: the start of each P's dependencies in the sweep's dependency list
// (cols[e] below: the particle that row P depends on); R: ring capacity in rows (a power of two, at least
// 3 * TP_ROWS_MAX; the kernel uses TP_RING).  Returns false when a row has
// more than TP_WAVE dependencies or R is unusable.
bool build_task_plan(int N, const int *lev, int nlev, const int *rowptr, int R, TaskPlan &tp);

// The lane arrays of a placed plan: src[e]: factor block of dependency e
// (nullptr: e); forward: the block is scaled by the consumer's diagonal block
// (forward sweep) or by the producer's (backward sweep).  Returns false if a
// dependency outside the ring is not at least three steps old (it cannot be,
// as three steps fit in the ring).
bool task_plan_lanes(int N, const int *rowptr, const int *cols, const int *src, bool forward, int R, TaskPlan &tp);
