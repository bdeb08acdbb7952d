// CPU emulation of tp_sweeps_kernel and its pre/post/pack kernels
// (fsd_amd/csrc/k_precond.hip) on the schedules of sweep_plan.cpp.
//
// The kernel's data flow is replayed step by step with a model of when each
// value becomes visible: a lane's operands for step S are fetched during step
// S - 1 (after the barrier that ends step S - 2); an owner's result enters the
// LDS ring during its step and global memory during the next step; every step
// ends with a workgroup barrier.  Checks: every index in bounds; every ring
// read finds the producer's row, written in an earlier step and not
// overwritten by a write of the same step; every global read finds the
// producer's lane, stored before the barrier the fetch follows; every
// reduction stays inside one wave and reads lanes of the same row; the
// result equals a direct block solve.
// usage: sweep_emulation N band max_lower_degree [isolated_percent] [R]
#include "../../fsd_amd/csrc/sweep_plan.h"
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <vector>

static long g_oob = 0, g_hazard = 0;
#define CHECK(cond, what)                                                  \
    do {                                                                   \
        if (!(cond)) {                                                     \
            if (g_hazard + g_oob < 10) printf("violation: %s\n", what);    \
            g_hazard++;                                                    \
        }                                                                  \
    } while (0)

static uint64_t tp_idx(uint64_t g, uint64_t i, uint64_t n) { return ((g >> 10) * n + i) * TP_LANES + (g & 1023); }

struct Lane { int P = -1, j = -1; };   // row and dependency index of a lane (j = -1: owner without block)

int main(int argc, char **argv)
{
    const int N = argc > 1 ? atoi(argv[1]) : 3000, band = argc > 2 ? atoi(argv[2]) : 40;
    const int maxdeg = argc > 3 ? atoi(argv[3]) : 5, iso = argc > 4 ? atoi(argv[4]) : 0;
    const int R = argc > 5 ? atoi(argv[5]) : (163840 / 16 - 3 * TP_LANES) / 3;
    std::mt19937 rng(5);
    std::vector<std::set<int>> adj(N);
    for (int P = 0; P < N; ++P) {
        if ((int)(rng() % 100) < iso) continue;
        const int deg = rng() % maxdeg;
        for (int t = 0; t < deg; ++t) {
            const int Q = P - 1 - (int)(rng() % band);
            if (Q >= 0) { adj[P].insert(Q); adj[Q].insert(P); }
        }
    }
    std::vector<int> lrow(N + 1), lcol, urow(N + 1, 0), ucol, upos;
    for (int P = 0; P < N; ++P) {
        lrow[P] = (int)lcol.size();
        for (int Q : adj[P]) if (Q < P) { lcol.push_back(Q); urow[Q + 1]++; }
    }
    lrow[N] = (int)lcol.size();
    for (int P = 0; P < N; ++P) urow[P + 1] += urow[P];
    ucol.resize(lcol.size());
    upos.resize(lcol.size());
    {
        std::vector<int> fill(urow.begin(), urow.end() - 1);
        for (int P = 0; P < N; ++P)
            for (int e = lrow[P]; e < lrow[P + 1]; ++e) { const int Q = lcol[e]; ucol[fill[Q]] = P; upos[fill[Q]] = e; fill[Q]++; }
    }
    const int64_t nlow = (int64_t)lcol.size();
    std::vector<int> lf(N, 0), lb(N, 0);
    int mf = 0, mb = 0;
    for (int P = 0; P < N; ++P) { int l = 0; for (int e = lrow[P]; e < lrow[P + 1]; ++e) l = std::max(l, lf[lcol[e]] + 1); lf[P] = l; mf = std::max(mf, l); }
    for (int P = N - 1; P >= 0; --P) { int l = 0; for (int e = urow[P]; e < urow[P + 1]; ++e) l = std::max(l, lb[ucol[e]] + 1); lb[P] = l; mb = std::max(mb, l); }
    TaskPlan tf, tb;
    if (!build_task_plan(N, lf.data(), mf + 1, lrow.data(), lcol.data(), nullptr, true, R, tf) ||
        !build_task_plan(N, lb.data(), mb + 1, urow.data(), ucol.data(), upos.data(), false, R, tb)) {
        printf("plan failed\n");
        return 1;
    }
    link_task_plans(N, tf, tb);

    // random factor (well conditioned), permutation and right-hand sides
    std::uniform_real_distribution<double> U(-0.3, 0.3);
    std::vector<double> Lblk(36 * nlow), Ldiag(36 * (size_t)N, 0.0);
    for (auto &v : Lblk) v = U(rng);
    for (int P = 0; P < N; ++P)
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c <= r; ++c) Ldiag[36 * P + 6 * r + c] = (r == c) ? 1.5 + U(rng) : U(rng);
    std::vector<int> perm(N);
    for (int i = 0; i < N; ++i) perm[i] = i;
    std::shuffle(perm.begin(), perm.end(), rng);
    std::vector<double> bvec[2];
    for (int r = 0; r < 2; ++r) { bvec[r].resize(6 * (size_t)N); for (auto &v : bvec[r]) v = U(rng); }
    // direct: y = L^-1 Pb, x = L^-T y, out[perm[P]] = x[P]
    std::vector<double> ref[2];
    for (int r = 0; r < 2; ++r) {
        std::vector<double> y(6 * (size_t)N), x(6 * (size_t)N);
        for (int P = 0; P < N; ++P) {
            double v[6];
            for (int a = 0; a < 6; ++a) v[a] = bvec[r][6 * perm[P] + a];
            for (int e = lrow[P]; e < lrow[P + 1]; ++e)
                for (int a = 0; a < 6; ++a) for (int c = 0; c < 6; ++c) v[a] -= Lblk[36 * e + 6 * a + c] * y[6 * lcol[e] + c];
            for (int a = 0; a < 6; ++a) { double s = v[a]; for (int c = 0; c < a; ++c) s -= Ldiag[36 * P + 6 * a + c] * y[6 * P + c]; y[6 * P + a] = s / Ldiag[36 * P + 7 * a]; }
        }
        for (int P = N - 1; P >= 0; --P) {
            double v[6];
            for (int a = 0; a < 6; ++a) v[a] = y[6 * P + a];
            for (int u = urow[P]; u < urow[P + 1]; ++u) {
                const int Q = ucol[u], e = upos[u];
                for (int a = 0; a < 6; ++a) for (int c = 0; c < 6; ++c) v[a] -= Lblk[36 * e + 6 * c + a] * x[6 * Q + c];
            }
            for (int a = 5; a >= 0; --a) { double s = v[a]; for (int c = a + 1; c < 6; ++c) s -= Ldiag[36 * P + 6 * c + a] * x[6 * P + c]; x[6 * P + a] = s / Ldiag[36 * P + 7 * a]; }
        }
        ref[r].resize(6 * (size_t)N);
        for (int P = 0; P < N; ++P) for (int a = 0; a < 6; ++a) ref[r][6 * perm[P] + a] = x[6 * P + a];
    }

    // inverse diagonal blocks (diag_inverse_kernel)
    std::vector<double> Z(36 * (size_t)N, 0.0);
    for (int P = 0; P < N; ++P) {
        const double *Lm = &Ldiag[36 * P];
        double *Zm = &Z[36 * P];
        for (int c = 0; c < 6; ++c) {
            Zm[7 * c] = 1.0 / Lm[7 * c];
            for (int r = c + 1; r < 6; ++r) { double s = 0; for (int q = c; q < r; ++q) s += Lm[6 * r + q] * Zm[6 * q + c]; Zm[6 * r + c] = -s / Lm[7 * r]; }
        }
    }
    struct Dev {
        const TaskPlan *tp;
        uint64_t lanes;
        std::vector<double> blk, inp, res;   // float4 slots x 4
        std::vector<Lane> lane;
    } D[2];
    const int *rowp[2] = {lrow.data(), urow.data()}, *colp[2] = {lcol.data(), ucol.data()};
    for (int sw = 0; sw < 2; ++sw) {
        Dev &d = D[sw];
        d.tp = sw ? &tb : &tf;
        d.lanes = (uint64_t)d.tp->nsteps * TP_LANES;
        d.blk.assign(d.lanes * 9 * 4, NAN);
        d.inp.assign(d.lanes * 3 * 4, NAN);
        d.res.assign(d.lanes * 3 * 4, NAN);
        d.lane.assign(d.lanes, Lane{});
        for (int P = 0; P < N; ++P) {
            const int g = d.tp->gtid[P], n = rowp[sw][P + 1] - rowp[sw][P];
            CHECK(g >= 0 && (uint64_t)g + std::max(n, 1) <= d.lanes, "owner lane out of range");
            d.lane[g] = Lane{P, n ? 0 : -1};
            for (int j = 1; j < n; ++j) d.lane[g + j] = Lane{P, j};
        }
        // tp_pack_kernel
        for (uint64_t g = 0; g < d.lanes; ++g) {
            const int e = d.tp->bsrc[2 * g], dg = d.tp->bsrc[2 * g + 1];
            if (e < 0) continue;
            CHECK(e < nlow && dg >= 0 && dg < N, "bsrc out of range");
            double M[36];
            for (int a = 0; a < 6; ++a)
                for (int c = 0; c < 6; ++c) {
                    double s = 0;
                    for (int q = 0; q <= a; ++q) s += Z[36 * (size_t)dg + 6 * a + q] * Lblk[36 * (size_t)e + 6 * q + c];
                    if (sw) M[6 * c + a] = s; else M[6 * a + c] = s;
                }
            for (int i = 0; i < 9; ++i) {
                const uint64_t k = tp_idx(g, i, 9);
                if (k >= d.lanes * 9) { g_oob++; continue; }
                for (int q = 0; q < 4; ++q) d.blk[4 * k + q] = M[4 * i + q];
            }
        }
    }
    // tp_pre_kernel
    for (int P = 0; P < N; ++P) {
        double w[12];
        for (int r = 0; r < 6; ++r) {
            double sa = 0, sb = 0;
            for (int c = 0; c <= r; ++c) { sa += Z[36 * P + 6 * r + c] * bvec[0][6 * perm[P] + c]; sb += Z[36 * P + 6 * r + c] * bvec[1][6 * perm[P] + c]; }
            w[r] = sa; w[6 + r] = sb;
        }
        for (int i = 0; i < 3; ++i) {
            const uint64_t k = tp_idx(tf.gtid[P], i, 3);
            if (k >= D[0].lanes * 3) { g_oob++; continue; }
            for (int q = 0; q < 4; ++q) D[0].inp[4 * k + q] = w[4 * i + q];
        }
    }

    // the sweeps
    long nring = 0, nglob = 0;
    std::vector<int> cissued(D[1].lanes, -1);   // step during which the copy into a backward input is issued
    for (int sw = 0; sw < 2; ++sw) {
        Dev &d = D[sw];
        const TaskPlan &tp = *d.tp;
        std::vector<double> ring(12 * (size_t)R, NAN), red(12 * TP_LANES, NAN);
        std::vector<int> ring_row(R, -1), ring_step(R, -1);
        std::vector<int> issued(d.lanes, -1);                  // step during which the lane's result store is issued
        auto lanevec = [&](const std::vector<double> &a, uint64_t g, double *w) {
            for (int i = 0; i < 3; ++i) {
                const uint64_t k = tp_idx(g, i, 3);
                if (k >= a.size() / 4) { g_oob++; return; }
                for (int q = 0; q < 4; ++q) w[4 * i + q] = a[4 * k + q];
            }
        };
        std::vector<std::pair<uint64_t, std::vector<double>>> pend;   // results of the previous step
        for (int S = 0; S < tp.nsteps; ++S) {
            // stores of the previous step, issued during this step
            for (auto &pw : pend) {
                const uint64_t g = pw.first;
                for (int i = 0; i < 3; ++i) for (int q = 0; q < 4; ++q) d.res[4 * tp_idx(g, i, 3) + q] = pw.second[4 * i + q];
                issued[g] = S;
                if (!sw) {
                    const int dest = tp.info[4 * g + 3];
                    if (dest < 0 || (uint64_t)dest >= D[1].lanes) { g_oob++; continue; }
                    for (int i = 0; i < 3; ++i) for (int q = 0; q < 4; ++q) D[1].inp[4 * tp_idx(dest, i, 3) + q] = pw.second[4 * i + q];
                    cissued[dest] = S;
                }
            }
            pend.clear();
            std::vector<double> part(12 * TP_LANES, 0.0);
            std::set<int> read_slots;
            std::vector<std::pair<int, std::vector<double>>> writes;   // ring writes of this step
            for (int t = 0; t < TP_LANES; ++t) {
                const uint64_t g = (uint64_t)S * TP_LANES + t;
                const int src = tp.info[4 * g], slot = tp.info[4 * g + 1];
                const Lane ln = d.lane[g];
                if (src >= 0) {
                    CHECK(ln.P >= 0 && ln.j >= 0, "block on a lane without dependency");
                    if (ln.P < 0 || ln.j < 0) continue;
                    const int Q = colp[sw][rowp[sw][ln.P] + ln.j];
                    double x[12];
                    if (src & TP_GLOBAL) {
                        const int h = src & ~TP_GLOBAL;
                        CHECK(h == tp.gtid[Q], "global read of the wrong lane");
                        // fetched during step S - 1, after the barrier ending step S - 2
                        CHECK(issued[h] >= 0 && issued[h] <= S - 2, "global read before the store is ordered");
                        lanevec(d.res, h, x);
                        ++nglob;
                    } else {
                        CHECK(src < R, "ring slot out of range");
                        if (src >= R) { g_oob++; continue; }
                        CHECK(ring_row[src] == Q && ring_step[src] >= 0 && ring_step[src] < S, "ring read of a wrong or unwritten row");
                        for (int q = 0; q < 12; ++q) x[q] = ring[12 * src + q];
                        read_slots.insert(src);
                        ++nring;
                    }
                    double M[36];
                    for (int i = 0; i < 9; ++i) for (int q = 0; q < 4; ++q) M[4 * i + q] = d.blk[4 * tp_idx(g, i, 9) + q];
                    for (int a = 0; a < 6; ++a)
                        for (int c = 0; c < 6; ++c) { part[12 * t + a] += M[6 * a + c] * x[c]; part[12 * t + 6 + a] += M[6 * a + c] * x[6 + c]; }
                }
                if (src >= 0 && slot < 0)
                    for (int q = 0; q < 12; ++q) red[12 * t + q] = part[12 * t + q];
            }
            for (int t = 0; t < TP_LANES; ++t) {
                const uint64_t g = (uint64_t)S * TP_LANES + t;
                const int src = tp.info[4 * g], slot = tp.info[4 * g + 1], nd = tp.info[4 * g + 2];
                if (slot < 0) continue;
                const Lane ln = d.lane[g];
                CHECK(ln.P >= 0 && ln.j <= 0, "owner lane does not own a row");
                if (ln.P < 0) continue;
                CHECK(nd == rowp[sw][ln.P + 1] - rowp[sw][ln.P], "owner's lane count");
                double in[12];
                if (sw) CHECK(cissued[g] >= 0, "backward input not written");   // ordered by the barrier between sweeps
                lanevec(d.inp, g, in);
                std::vector<double> w(12);
                for (int q = 0; q < 12; ++q) w[q] = in[q] - (src >= 0 ? part[12 * t + q] : 0.0);
                for (int dd = 1; dd < nd; ++dd) {
                    const int u = t + dd;
                    CHECK(u < TP_LANES && u / TP_WAVE == t / TP_WAVE, "reduction leaves the wave");
                    if (u >= TP_LANES) { g_oob++; continue; }
                    const Lane lu = d.lane[(uint64_t)S * TP_LANES + u];
                    CHECK(lu.P == ln.P && lu.j == dd && tp.info[4 * ((uint64_t)S * TP_LANES + u) + 1] < 0, "reduction reads a lane of another row");
                    for (int q = 0; q < 12; ++q) w[q] -= red[12 * u + q];
                }
                CHECK(slot < R, "owner slot out of range");
                writes.push_back({slot, w});
                pend.push_back({g, w});
                ring_step[slot] = S;
                ring_row[slot] = ln.P;
            }
            for (auto &wr : writes) {
                CHECK(!read_slots.count(wr.first), "ring slot written and read in the same step");
                for (int q = 0; q < 12; ++q) ring[12 * wr.first + q] = wr.second[q];
            }
        }
        // stores of the last step after the loop, then the barrier between the sweeps / kernel end
        for (auto &pw : pend) {
            const uint64_t g = pw.first;
            for (int i = 0; i < 3; ++i) for (int q = 0; q < 4; ++q) d.res[4 * tp_idx(g, i, 3) + q] = pw.second[4 * i + q];
            if (!sw) {
                const int dest = tp.info[4 * g + 3];
                if (dest < 0 || (uint64_t)dest >= D[1].lanes) { g_oob++; continue; }
                for (int i = 0; i < 3; ++i) for (int q = 0; q < 4; ++q) D[1].inp[4 * tp_idx(dest, i, 3) + q] = pw.second[4 * i + q];
                cissued[dest] = tp.nsteps;
            }
        }
        if (!sw)   // every backward input is written before the barrier between the sweeps
            for (uint64_t g = 0; g < D[1].lanes; ++g)
                if (tb.info[4 * g + 1] >= 0) CHECK(cissued[g] >= 0, "backward input never written");
    }
    // tp_post_kernel
    std::vector<double> out[2];
    out[0].assign(6 * (size_t)N, NAN);
    out[1].assign(6 * (size_t)N, NAN);
    for (int P = 0; P < N; ++P) {
        double u[12];
        for (int i = 0; i < 3; ++i) for (int q = 0; q < 4; ++q) u[4 * i + q] = D[1].res[4 * tp_idx(tb.gtid[P], i, 3) + q];
        for (int a = 0; a < 6; ++a) {
            double sa = 0, sb = 0;
            for (int c = a; c < 6; ++c) { sa += Z[36 * P + 6 * c + a] * u[c]; sb += Z[36 * P + 6 * c + a] * u[6 + c]; }
            out[0][6 * perm[P] + a] = sa;
            out[1][6 * perm[P] + a] = sb;
        }
    }
    double err = 0, mx = 0;
    for (size_t i = 0; i < 6 * (size_t)N; ++i) {
        err = std::max(err, std::fabs(out[0][i] - ref[0][i]));
        err = std::max(err, std::fabs(out[1][i] - ref[1][i]));
        mx = std::max(mx, std::fabs(ref[0][i]));
    }
    if (!(err < 1e-9 * std::max(1.0, mx))) printf("result mismatch\n");
    printf("N=%d nlow=%ld steps f=%d b=%d ring reads=%ld global reads=%ld hazards=%ld oob=%ld max err=%.3e (max |x|=%.3e)\n",
           N, (long)nlow, tf.nsteps, tb.nsteps, nring, nglob, g_hazard, g_oob, err, mx);
    return (g_hazard || g_oob || !(err < 1e-9 * std::max(1.0, mx))) ? 1 : 0;
}
