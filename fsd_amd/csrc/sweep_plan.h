// Schedules of the block-triangular sweeps (host side).
#pragma once
#include <cstdint>
#include <vector>

struct SweepPlan {
    int nlev = 0;
    std::vector<int> levrow;   // nlev+1: first position of each level
    std::vector<int> levjd;    // nlev+1: first record of each level
    std::vector<int> rec;      // 2 per record: (first entry, number of rows with more than j blocks)
    std::vector<int> order;    // position -> permuted particle
    std::vector<int> pos;      // permuted particle -> position
    std::vector<int> esrc;     // entry -> index of its block in the factor's block array
    std::vector<int> ecol;     // entry -> position (in this sweep) of the row it depends on
};

// lev[P] in [0, nlev): level of permuted particle P in the sweep's DAG;
// rowptr/cols: the sweep's dependencies of each P (cols[e] = particle the
// row P depends on); src[e]: block index of dependency e (nullptr: e itself).
void build_sweep_plan(int N, const int *lev, int nlev, const int *rowptr, const int *cols, const int *src,
                      SweepPlan &pl);
