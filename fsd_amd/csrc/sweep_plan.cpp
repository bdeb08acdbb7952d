// Static schedules of the two block-triangular sweeps (see sweep_plan.h).
//
// Rows are taken level by level; inside a level by descending number of
// dependencies (ties by particle id; one counting sort) and packed first-fit
// into the waves of the current step, so busy waves are the low-numbered ones
// and spread over the four SIMDs.  A level that does not fit in 1024 lanes or TP_ROWS_MAX rows
// continues in the next step (its rows are independent, so the extra barrier
// is harmless).  Sequence numbers follow the packing order, so a wave's
// owners hold nearly consecutive numbers and their per-row loads and stores
// coalesce.
#include "sweep_plan.h"

#include <algorithm>

bool build_task_plan(int N, const int *lev, int nlev, const int *rowptr, int R, TaskPlan &tp)
{
    tp.nsteps = 0;   // the vectors keep their capacity from the previous build (no page faults)
    if (R < 3 * TP_ROWS_MAX || (R & (R - 1)) != 0 || N >= (1 << TP_ND_SHIFT)) return false;
    auto cnt = [&](int P) { return rowptr[P + 1] - rowptr[P]; };
    for (int P = 0; P < N; ++P)
        if (cnt(P) > TP_WAVE) return false;
    // rows by level, then by descending count, then by particle: one counting sort on
    // the key level * (TP_WAVE + 1) + (TP_WAVE - count)
    const size_t nkey = (size_t)nlev * (TP_WAVE + 1);
    std::vector<int> &kptr = tp.kptr, &order = tp.order, &lptr = tp.lptr;
    kptr.assign(nkey + 1, 0);
    order.resize(N);
    lptr.assign(nlev + 1, 0);
    for (int P = 0; P < N; ++P) kptr[(size_t)lev[P] * (TP_WAVE + 1) + (TP_WAVE - cnt(P)) + 1]++;
    for (size_t k = 0; k < nkey; ++k) kptr[k + 1] += kptr[k];
    for (int l = 0; l <= nlev; ++l) lptr[l] = kptr[(size_t)l * (TP_WAVE + 1)];
    for (int P = 0; P < N; ++P) order[kptr[(size_t)lev[P] * (TP_WAVE + 1) + (TP_WAVE - cnt(P))]++] = P;

    // packing: lanes, sequence numbers and steps
    constexpr int NWAVE = TP_LANES / TP_WAVE;
    std::vector<int> &gtid = tp.lane, &rowstep = tp.step, &send = tp.send;
    gtid.resize(N);
    rowstep.resize(N);
    send.clear();
    tp.seq.resize(N);
    int nseq = 0, step = -1, rows = 0;
    int fill[NWAVE];
    auto open_step = [&]() {
        ++step;
        rows = 0;
        std::fill(fill, fill + NWAVE, 0);
    };
    int wrows[NWAVE][TP_WAVE], wn[NWAVE] = {};   // rows of each wave in the open step (one lane at least each)
    auto flush = [&]() {
        // sequence numbers wave by wave, in lane order inside a wave
        for (int w = 0; w < NWAVE; ++w) {
            for (int m = 0; m < wn[w]; ++m) tp.seq[wrows[w][m]] = nseq++;
            wn[w] = 0;
        }
        send.push_back(nseq);
    };
    for (int l = 0; l < nlev; ++l) {
        if (lptr[l + 1] == lptr[l]) continue;
        open_step();
        for (int k = lptr[l]; k < lptr[l + 1]; ++k) {
            const int P = order[k];
            const int need = std::max(cnt(P), 1);
            int w = 0;
            while (w < NWAVE && fill[w] + need > TP_WAVE) ++w;
            if (w == NWAVE || rows == TP_ROWS_MAX) {
                flush();
                open_step();
                w = 0;
            }
            gtid[P] = step * TP_LANES + w * TP_WAVE + fill[w];
            wrows[w][wn[w]++] = P;
            fill[w] += need;
            ++rows;
            rowstep[P] = step;
        }
        flush();
    }
    tp.nsteps = step + 1;
    return true;
}

bool task_plan_lanes(int N, const int *rowptr, const int *cols, const int *src, bool forward, int R, TaskPlan &tp)
{
    const std::vector<int> &send = tp.send;
    auto in_ring = [&](int q, int S) { return q >= send[S] - R; };
    const size_t L = (size_t)tp.nsteps * TP_LANES;
    tp.info.assign(2 * L, -1);   // reuses the capacity of the previous build
    tp.bsrc.assign(2 * L, -1);
    for (int P = 0; P < N; ++P) {
        const int g = tp.lane[P], S = tp.step[P], n = rowptr[P + 1] - rowptr[P];
        tp.info[2 * (size_t)g + 1] = tp.seq[P] | (n << TP_ND_SHIFT);
        for (int j = 0; j < n; ++j) {
            const int e = rowptr[P] + j, Q = cols[e];
            const size_t h = (size_t)g + j;
            if (in_ring(tp.seq[Q], S)) {
                tp.info[2 * h] = tp.seq[Q] & (R - 1);
            } else {
                if (tp.step[Q] > S - 3) return false;
                tp.info[2 * h] = TP_GLOBAL | tp.seq[Q];
            }
            tp.bsrc[2 * h] = src ? src[e] : e;
            tp.bsrc[2 * h + 1] = forward ? P : Q;
        }
    }
    return true;
}
