// Host phase of Precondition_Wrap (Precondition.cu:1326-1497): the reordering
// of Precondition_Reorder (RCM on the pruned graph), the block pattern of the
// permuted matrix and the level schedules of its triangular sweeps.
#include "precond_plan.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <thread>

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// fn(p0, p1) over `threads` contiguous ranges of [0, N)
template <class F>
void ranges(int N, int threads, F &&fn)
{
    std::vector<std::thread> th;
    for (int t = 1; t < threads; ++t)
        th.emplace_back(fn, (int)((int64_t)N * t / threads), (int)((int64_t)N * (t + 1) / threads));
    fn(0, (int)((int64_t)N / threads));
    for (auto &x : th) x.join();
}

// level of every row of a DAG given by its dependencies (rows in dependency order
// when `ascending`, else in reverse), and the rows bucketed by level
void levels(int N, const std::vector<int> &row, const std::vector<int> &col, bool ascending, std::vector<int> &lev,
            int &nlev, std::vector<int> &ptr, std::vector<int> &items)
{
    lev.assign(N, 0);
    int mx = 0;
    for (int k = 0; k < N; ++k) {
        const int P = ascending ? k : N - 1 - k;
        int l = 0;
        for (int e = row[P]; e < row[P + 1]; ++e) l = std::max(l, lev[col[e]] + 1);
        lev[P] = l;
        mx = std::max(mx, l);
    }
    nlev = N ? mx + 1 : 0;
    ptr.assign(nlev + 1, 0);
    for (int P = 0; P < N; ++P) ptr[lev[P] + 1]++;
    for (int l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
    items.resize(N);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int P = 0; P < N; ++P) items[fill[lev[P]]++] = P;
}

}  // namespace

void plan_precond(int N, const int *pn, const int *plstride, int stride, int ring, int threads, PrecondHost &h,
                  bool keep_order)
{
    const double t0 = now_ms();
    threads = std::max(1, threads);
    // RCM of the pruned graph (self loops included), as Precondition_Reorder passes it to genrcm;
    // keep_order keeps the last plan's ordering instead (any ordering gives a valid factor)
    if (!keep_order || (int)h.perm.size() != N || (int)h.iperm.size() != N) {
        h.perm.resize(N);
        host_rcm_rows(N, pn, plstride, stride, h.perm.data());
        h.iperm.resize(N);
        for (int P = 0; P < N; ++P) h.iperm[h.perm[P]] = P;
    }
    const double t1 = now_ms();

    // block adjacency on the permuted graph.  The graph is symmetric, so row P's blocks
    // above the diagonal are its neighbours Q > P.  One parallel pass writes each row's
    // permuted neighbours sorted (self included, padded with INT_MAX) into a 16-wide
    // scratch row; a second copies the lower and upper parts into CSR form and finds, for
    // each upper block (P, Q), the position of P among Q's lower blocks.  Ranks are counted
    // with comparisons rather than branches (the rows are short and their order random).
    const int *perm = h.perm.data(), *iperm = h.iperm.data();
    constexpr int K = PLAN_MAX_ROW;
    h.srt.resize((size_t)N * K);
    h.nlo.resize(N);
    const int nthr = N >= 8192 ? threads : 1;
    ranges(N, nthr, [&](int p0, int p1) {
        for (int P = p0; P < p1; ++P) {
            // rows are visited in random particle order: prefetch the list 16 rows ahead and
            // the permuted ids of the list 8 rows ahead (fetched by then)
            if (P + 16 < p1) __builtin_prefetch(plstride + (size_t)perm[P + 16] * stride);
            if (P + 8 < p1) {
                const int i8 = perm[P + 8];
                const int *r8 = plstride + (size_t)i8 * stride;
                for (int s = 0; s < pn[i8]; ++s) __builtin_prefetch(iperm + r8[s]);
            }
            const int i = perm[P], n = pn[i];
            const int *r = plstride + (size_t)i * stride;
            int q[K];
            for (int s = 0; s < n; ++s) q[s] = iperm[r[s]];
            int *srt = h.srt.data() + (size_t)P * K;
            for (int s = 0; s < K; ++s) srt[s] = INT_MAX;
            int nl = 0;
            for (int s = 0; s < n; ++s) {
                int rank = 0;
                for (int t = 0; t < n; ++t) rank += q[t] < q[s];
                srt[rank] = q[s];
                nl += q[s] < P;
            }
            h.nlo[P] = nl;
        }
    });
    h.lrow.resize(N + 1);
    h.urow.resize(N + 1);
    h.lrow[0] = h.urow[0] = 0;
    int64_t nnz_ref = 0;
    for (int P = 0; P < N; ++P) {
        const int n = pn[perm[P]];
        h.lrow[P + 1] = h.lrow[P] + h.nlo[P];
        h.urow[P + 1] = h.urow[P] + (n - 1 - h.nlo[P]);
        nnz_ref += (n > 1) ? 36 * (int64_t)n : 6;
    }
    h.lcol.resize(h.lrow[N]);
    h.ucol.resize(h.urow[N]);
    h.upos.resize(h.urow[N]);
    ranges(N, nthr, [&](int p0, int p1) {
        for (int P = p0; P < p1; ++P) {
            if (P + 8 < p1) {   // the scratch rows of the upper neighbours 8 rows ahead
                const int *s8 = h.srt.data() + (size_t)(P + 8) * K;
                for (int u = h.nlo[P + 8] + 1; u < K && s8[u] != INT_MAX; ++u)
                    __builtin_prefetch(h.srt.data() + (size_t)s8[u] * K);
            }
            const int *srt = h.srt.data() + (size_t)P * K;
            const int nl = h.nlo[P], nu = h.urow[P + 1] - h.urow[P];
            std::copy(srt, srt + nl, h.lcol.begin() + h.lrow[P]);
            for (int u = 0; u < nu; ++u) {
                const int Q = srt[nl + 1 + u];            // block (Q, P) below the diagonal
                const int *sq = h.srt.data() + (size_t)Q * K;
                int k = 0;
                for (int t = 0; t < K; ++t) k += sq[t] < P;   // P's place among Q's lower neighbours
                h.ucol[h.urow[P] + u] = Q;
                h.upos[h.urow[P] + u] = h.lrow[Q] + k;
            }
        }
    });
    h.nlow = (int64_t)h.lcol.size();
    h.nnz_ref = nnz_ref;
    const double t2 = now_ms();

    // level schedules (forward level = 1 + max level of the lower neighbours, backward
    // level = 1 + max level of the upper neighbours, counted from N-1) and the placement
    // of the rows of the single-workgroup sweeps: the two directions are independent, the
    // backward one runs on a second thread
    bool okf = false, okb = false;
    auto direction = [&](bool fwd) {
        std::vector<int> &lev = fwd ? h.lev_f : h.lev_b;
        const std::vector<int> &row = fwd ? h.lrow : h.urow, &col = fwd ? h.lcol : h.ucol;
        int &nlev = fwd ? h.nlev_f : h.nlev_b;
        levels(N, row, col, fwd, lev, nlev, fwd ? h.lev_f_ptr : h.lev_b_ptr, fwd ? h.itf : h.itb);
        if (ring > 0) (fwd ? okf : okb) = build_task_plan(N, lev.data(), nlev, row.data(), ring, fwd ? h.tp_f : h.tp_b);
    };
    {
        std::thread tb(direction, false);
        direction(true);
        tb.join();
    }
    h.tp_ok = ring > 0 && okf && okb;
    const double t3 = now_ms();
    h.ms_rcm = t1 - t0;
    h.ms_adj = t2 - t1;
    h.ms_lev = t3 - t2;
}
