// Static schedules of the two block-triangular sweeps of G = P^T L^-T L^-1 P.
//
// Rows (permuted particles) are grouped by level of the sweep's DAG; inside a
// level they are ordered by descending number of off-diagonal blocks (ties by
// particle id), and the blocks of the level are stored in jagged-diagonal
// order: first the 0th block of every row that has one, then the 1st, ... .
// A level's rows are therefore consecutive positions r0..r1-1 and the j-th
// block of row r0+k sits at entry rec[j].first + k, so a solver thread finds
// its blocks from two numbers per (level, j) with no per-row pointer loads.
#include "sweep_plan.h"

#include <algorithm>

void build_sweep_plan(int N, const int *lev, int nlev, const int *rowptr, const int *cols, const int *src,
                      SweepPlan &pl)
{
    pl.nlev = nlev;
    pl.levrow.assign(nlev + 1, 0);
    for (int P = 0; P < N; ++P) pl.levrow[lev[P] + 1]++;
    for (int l = 0; l < nlev; ++l) pl.levrow[l + 1] += pl.levrow[l];
    pl.order.assign(N, 0);
    {
        std::vector<int> fill(pl.levrow.begin(), pl.levrow.end() - 1);
        for (int P = 0; P < N; ++P) pl.order[fill[lev[P]]++] = P;   // ascending P inside a level
    }
    auto cnt = [&](int P) { return rowptr[P + 1] - rowptr[P]; };
    for (int l = 0; l < nlev; ++l)
        std::stable_sort(pl.order.begin() + pl.levrow[l], pl.order.begin() + pl.levrow[l + 1],
                         [&](int a, int b) { return cnt(a) > cnt(b); });
    pl.pos.assign(N, 0);
    for (int k = 0; k < N; ++k) pl.pos[pl.order[k]] = k;
    pl.levjd.assign(nlev + 1, 0);
    pl.rec.clear();
    pl.esrc.clear();
    pl.ecol.clear();
    const int64_t nent = rowptr[N];
    pl.esrc.reserve(nent);
    pl.ecol.reserve(nent);
    for (int l = 0; l < nlev; ++l) {
        pl.levjd[l] = (int)(pl.rec.size() / 2);
        const int r0 = pl.levrow[l], r1 = pl.levrow[l + 1];
        const int J = (r1 > r0) ? cnt(pl.order[r0]) : 0;   // rows sorted by descending count
        for (int j = 0; j < J; ++j) {
            int nr = 0;
            while (r0 + nr < r1 && cnt(pl.order[r0 + nr]) > j) ++nr;
            pl.rec.push_back((int)pl.esrc.size());
            pl.rec.push_back(nr);
            for (int k = 0; k < nr; ++k) {
                const int P = pl.order[r0 + k];
                const int e = rowptr[P] + j;
                pl.esrc.push_back(src ? src[e] : e);
                pl.ecol.push_back(cols[e]);              // particle id for now, position below
            }
        }
    }
    pl.levjd[nlev] = (int)(pl.rec.size() / 2);
    for (auto &c : pl.ecol) c = pl.pos[c];
}
