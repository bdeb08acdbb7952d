// Host phase of the preconditioner build (host only, no HIP): the RCM order of
// the pruned particle graph, the block adjacency below and above the diagonal
// on the permuted graph, the level schedules of both triangular sweeps and
// their single-workgroup plans.
#pragma once
#include <cstdint>
#include <vector>

#include "sweep_plan.h"

// host RCM (George & Liu GENRCM) of a symmetric graph with self loops: packed
// (phead[N+1], plist) or fixed-stride (cnt[v] neighbours of v at rows[v * stride])
void host_rcm(int N, const int *phead, const int *plist, int *perm);
void host_rcm_rows(int N, const int *cnt, const int *rows, int stride, int *perm);

// entries of a pruned list at most (self included)
constexpr int PLAN_MAX_ROW = 16;

struct PrecondHost {
    std::vector<int> perm, iperm;              // permuted index -> particle, particle -> permuted index
    std::vector<int> lrow, lcol;               // lower blocks (Q < P) of each permuted row P, ascending
    std::vector<int> urow, ucol, upos;         // upper blocks: rows P > Q of column Q and their lower index
    std::vector<int> lev_f_ptr, lev_b_ptr;     // level starts of the forward / backward DAG
    std::vector<int> itf, itb;                 // permuted particles by forward / backward level
    std::vector<int> lev_f, lev_b;             // level of every permuted particle
    int nlev_f = 0, nlev_b = 0;
    int64_t nlow = 0;                          // lower blocks
    int64_t nnz_ref = 0;                       // nonzeros of the reference's scalar CSR
    std::vector<int> srt, nlo;                 // scratch: sorted permuted neighbours (PLAN_MAX_ROW per row), lower count
    TaskPlan tp_f, tp_b;                       // single-workgroup sweep plans (tp_ok: both built)
    bool tp_ok = false;
    double ms_rcm = 0, ms_adj = 0, ms_lev = 0;   // host wall time of the last plan's phases (ms_lev: levels and sweep plans)
};

// pn[i] <= PLAN_MAX_ROW: entries of particle i's pruned list (self included,
// ascending), stored at plist[i * stride ...]; ring: LDS ring rows of the sweep kernel (0: no
// single-workgroup plans); threads: worker threads for the parallel parts; keep_order: keep the
// ordering of the last plan in h (of the same N) instead of computing the RCM of this graph.
void plan_precond(int N, const int *pn, const int *plist, int stride, int ring, int threads, PrecondHost &h,
                  bool keep_order = false);
