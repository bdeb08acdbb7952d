// CPU emulation of tp_sweep_kernel and its pre/mid/post/pack kernels
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
// usage: sweep_emulation N band max_lower_degree [isolated_percent] [R (power of two)]
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
    const int R = argc > 5 ? atoi(argv[5]) : TP_RING;
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
    if (!build_task_plan(N, lf.data(), mf + 1, lrow.data(), R, tf) ||
        !build_task_plan(N, lb.data(), mb + 1, urow.data(), R, tb) ||
        !task_plan_lanes(N, lrow.data(), lcol.data(), nullptr, true, R, tf) ||
        !task_plan_lanes(N, urow.data(), ucol.data(), upos.data(), false, R, tb)) {
        printf("plan failed\n");
        return 1;
    }

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
        std::vector<double> blk;        // float4 slots x 4, [step][9][1024]
        std::vector<double> inp, res;   // [3][N] float4 -> 12 doubles per row, slot i at 4 (i N + k)
        std::vector<Lane> lane;
    } D[2];
    const int *rowp[2] = {lrow.data(), urow.data()}, *colp[2] = {lcol.data(), ucol.data()};
    const int KM = (1 << TP_ND_SHIFT) - 1;
    for (int sw = 0; sw < 2; ++sw) {
        Dev &d = D[sw];
        d.tp = sw ? &tb : &tf;
        d.lanes = (uint64_t)d.tp->nsteps * TP_LANES;
        d.blk.assign(d.lanes * 9 * 4, NAN);
        d.inp.assign(12 * (size_t)N, NAN);
        d.res.assign(12 * (size_t)N, NAN);
        d.lane.assign(d.lanes, Lane{});
        std::vector<int> seen(N, 0), inv(N, -1);
        for (int Q = 0; Q < N; ++Q) {
            const int k = d.tp->seq[Q];
            CHECK(k >= 0 && k < N && inv[k] < 0, "sequence numbers are not a permutation");
            if (k >= 0 && k < N) inv[k] = Q;
        }
        for (uint64_t g = 0; g < d.lanes; ++g) {
            const int own = d.tp->info[2 * g + 1];
            if (own < 0) continue;
            const int k = own & KM, nd = own >> TP_ND_SHIFT;
            CHECK(k >= 0 && k < N, "sequence number out of range");
            if (k < 0 || k >= N) continue;
            const int P = inv[k];
            CHECK(P >= 0 && !seen[P], "row owned twice");
            if (P < 0) continue;
            seen[P] = 1;
            CHECK(nd == rowp[sw][P + 1] - rowp[sw][P], "owner's lane count");
            CHECK(g + std::max(nd, 1) <= d.lanes, "row lanes out of range");
            d.lane[g] = Lane{P, nd ? 0 : -1};
            for (int j = 1; j < nd; ++j) d.lane[g + j] = Lane{P, j};
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
    auto rowget = [&](const std::vector<double> &a, int k, double *w) {
        for (int i = 0; i < 3; ++i) for (int q = 0; q < 4; ++q) w[4 * i + q] = a[4 * ((size_t)i * N + k) + q];
    };
    auto rowput = [&](std::vector<double> &a, int k, const double *w) {
        for (int i = 0; i < 3; ++i) for (int q = 0; q < 4; ++q) a[4 * ((size_t)i * N + k) + q] = w[4 * i + q];
    };
    // tp_pre_kernel
    for (int P = 0; P < N; ++P) {
        double w[12];
        for (int r = 0; r < 6; ++r) {
            double sa = 0, sb = 0;
            for (int c = 0; c <= r; ++c) { sa += Z[36 * P + 6 * r + c] * bvec[0][6 * perm[P] + c]; sb += Z[36 * P + 6 * r + c] * bvec[1][6 * perm[P] + c]; }
            w[r] = sa; w[6 + r] = sb;
        }
        rowput(D[0].inp, tf.seq[P], w);
    }

    // the sweeps (one launch each), with tp_mid_kernel in between
    long nring = 0, nglob = 0;
    for (int sw = 0; sw < 2; ++sw) {
        Dev &d = D[sw];
        const TaskPlan &tp = *d.tp;
        if (sw) {   // tp_mid_kernel
            for (int P = 0; P < N; ++P) { double w[12]; rowget(D[0].res, tf.seq[P], w); rowput(D[1].inp, tb.seq[P], w); }
        }
        std::vector<double> ring(12 * (size_t)R, NAN), red(12 * TP_LANES, NAN);
        std::vector<int> ring_row(R, -1), ring_step(R, -1);
        std::vector<int> issued(N, -1);   // step during which the store of result k is issued
        std::vector<std::pair<int, std::vector<double>>> pend;   // results of the previous step
        for (int S = 0; S < tp.nsteps; ++S) {
            for (auto &pw : pend) { rowput(d.res, pw.first, pw.second.data()); issued[pw.first] = S; }
            pend.clear();
            std::vector<double> part(12 * TP_LANES, 0.0);
            std::set<int> read_slots;
            std::vector<std::pair<int, std::vector<double>>> writes;
            for (int t = 0; t < TP_LANES; ++t) {
                const uint64_t g = (uint64_t)S * TP_LANES + t;
                const int src = tp.info[2 * g], own = tp.info[2 * g + 1];
                const Lane ln = d.lane[g];
                if (src >= 0) {
                    CHECK(ln.P >= 0 && ln.j >= 0, "block on a lane without dependency");
                    if (ln.P < 0 || ln.j < 0) continue;
                    const int Q = colp[sw][rowp[sw][ln.P] + ln.j];
                    double x[12];
                    if (src & TP_GLOBAL) {
                        const int q = src & ~TP_GLOBAL;
                        CHECK(q == tp.seq[Q], "global read of the wrong row");
                        // fetched during step S - 1, after the barrier ending step S - 2
                        CHECK(issued[q] >= 0 && issued[q] <= S - 2, "global read before the store is ordered");
                        rowget(d.res, q, x);
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
                if (src >= 0 && own < 0)
                    for (int q = 0; q < 12; ++q) red[12 * t + q] = part[12 * t + q];
            }
            for (int t = 0; t < TP_LANES; ++t) {
                const uint64_t g = (uint64_t)S * TP_LANES + t;
                const int src = tp.info[2 * g], own = tp.info[2 * g + 1];
                if (own < 0) continue;
                const Lane ln = d.lane[g];
                if (ln.P < 0) continue;
                const int k = own & KM, nd = own >> TP_ND_SHIFT;
                double in[12];
                rowget(d.inp, k, in);
                std::vector<double> w(12);
                for (int q = 0; q < 12; ++q) w[q] = in[q] - (src >= 0 ? part[12 * t + q] : 0.0);
                for (int dd = 1; dd < nd; ++dd) {
                    const int u = t + dd;
                    CHECK(u < TP_LANES && u / TP_WAVE == t / TP_WAVE, "reduction leaves the wave");
                    if (u >= TP_LANES) { g_oob++; continue; }
                    const uint64_t gu = (uint64_t)S * TP_LANES + u;
                    const Lane lu = d.lane[gu];
                    CHECK(lu.P == ln.P && lu.j == dd && tp.info[2 * gu + 1] < 0, "reduction reads a lane of another row");
                    for (int q = 0; q < 12; ++q) w[q] -= red[12 * u + q];
                }
                const int slot = k & (R - 1);
                writes.push_back({slot, w});
                pend.push_back({k, w});
                ring_step[slot] = S;
                ring_row[slot] = ln.P;
            }
            for (auto &wr : writes) {
                CHECK(!read_slots.count(wr.first), "ring slot written and read in the same step");
                for (int q = 0; q < 12; ++q) ring[12 * wr.first + q] = wr.second[q];
            }
        }
        for (auto &pw : pend) rowput(d.res, pw.first, pw.second.data());   // after the loop; kernel end orders them
    }
    // tp_post_kernel
    std::vector<double> out[2];
    out[0].assign(6 * (size_t)N, NAN);
    out[1].assign(6 * (size_t)N, NAN);
    for (int P = 0; P < N; ++P) {
        double u[12];
        rowget(D[1].res, tb.seq[P], u);
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
