// CPU check of plan_precond (fsd_amd/csrc/precond_plan.cpp) with and without the kept ordering:
// a random geometric graph in the pruned-list layout (self first, at most 15 neighbours), planned
// once, then perturbed (pairs removed and added) and planned again with keep_order.  Checks:
// the ordering is the first plan's; the block pattern is the new graph's under it (every edge
// once, lower columns below the row and ascending, upper columns above it, upos pointing at the
// mirror lower block); the level schedules order every dependency; without keep_order the
// ordering is the RCM of the new graph.
// usage: plan_keep_order N seed
#include "../../fsd_amd/csrc/precond_plan.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <utility>
#include <vector>

static int g_bad = 0;
#define CHECK(cond, what)                                        \
    do {                                                         \
        if (!(cond)) {                                           \
            if (g_bad < 10) printf("violation: %s\n", what);     \
            g_bad++;                                             \
        }                                                        \
    } while (0)

using Edges = std::set<std::pair<int, int>>;   // (i, j), i < j

static void to_lists(int N, const Edges &E, std::vector<int> &pn, std::vector<int> &pl)
{
    pn.assign(N, 1);
    pl.assign((size_t)N * 16, 0);
    for (int i = 0; i < N; ++i) pl[(size_t)i * 16] = i;
    for (auto [i, j] : E) {
        pl[(size_t)i * 16 + pn[i]++] = j;
        pl[(size_t)j * 16 + pn[j]++] = i;
    }
}

static void check_pattern(int N, const Edges &E, const PrecondHost &h)
{
    CHECK((int64_t)E.size() == h.nlow, "one lower block per edge");
    Edges seen;
    for (int P = 0; P < N; ++P) {
        for (int e = h.lrow[P]; e < h.lrow[P + 1]; ++e) {
            const int Q = h.lcol[e];
            CHECK(Q < P, "lower column below the row");
            if (e > h.lrow[P]) CHECK(h.lcol[e - 1] < Q, "lower columns ascending");
            const int a = h.perm[P], b = h.perm[Q];
            seen.insert({std::min(a, b), std::max(a, b)});
            CHECK(h.lev_f[Q] < h.lev_f[P], "forward level after its dependencies");
        }
        for (int u = h.urow[P]; u < h.urow[P + 1]; ++u) {
            const int Q = h.ucol[u];
            CHECK(Q > P, "upper column above the row");
            CHECK(h.upos[u] >= h.lrow[Q] && h.upos[u] < h.lrow[Q + 1] && h.lcol[h.upos[u]] == P,
                  "upos at the mirror lower block");
            CHECK(h.lev_b[Q] < h.lev_b[P], "backward level after its dependencies");
        }
    }
    CHECK(seen == E, "the pattern is the graph's");
}

int main(int argc, char **argv)
{
    const int N = argc > 1 ? atoi(argv[1]) : 2000;
    std::mt19937 rng(argc > 2 ? atoi(argv[2]) : 1);
    // random geometric graph in a unit cube, degree capped at 15
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> x(N), y(N), z(N);
    for (int i = 0; i < N; ++i) x[i] = U(rng), y[i] = U(rng), z[i] = U(rng);
    const double rc = 1.2 / std::cbrt((double)N);
    Edges E;
    std::vector<int> deg(N, 0);
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) {
            const double d2 = (x[i] - x[j]) * (x[i] - x[j]) + (y[i] - y[j]) * (y[i] - y[j]) + (z[i] - z[j]) * (z[i] - z[j]);
            if (d2 < rc * rc && deg[i] < 15 && deg[j] < 15) E.insert({i, j}), deg[i]++, deg[j]++;
        }
    std::vector<int> pn, pl;
    to_lists(N, E, pn, pl);
    PrecondHost h;
    plan_precond(N, pn.data(), pl.data(), 16, 0, 2, h);
    check_pattern(N, E, h);
    const std::vector<int> perm0 = h.perm;
    // perturb: drop every 10th edge, add pairs of close nodes with room
    Edges E2;
    int k = 0;
    std::fill(deg.begin(), deg.end(), 0);
    for (auto e : E)
        if (k++ % 10) E2.insert(e), deg[e.first]++, deg[e.second]++;
    for (int t = 0; t < N; ++t) {
        const int i = rng() % N, j = rng() % N;
        if (i != j && deg[i] < 15 && deg[j] < 15 && E2.insert({std::min(i, j), std::max(i, j)}).second) deg[i]++, deg[j]++;
    }
    to_lists(N, E2, pn, pl);
    plan_precond(N, pn.data(), pl.data(), 16, 0, 2, h, true);
    CHECK(h.perm == perm0, "kept ordering");
    check_pattern(N, E2, h);
    plan_precond(N, pn.data(), pl.data(), 16, 0, 2, h, false);
    std::vector<int> rcm(N);
    host_rcm_rows(N, pn.data(), pl.data(), 16, rcm.data());
    CHECK(h.perm == rcm, "fresh ordering is the new graph's RCM");
    CHECK(h.perm != perm0, "the perturbed graph orders differently");
    check_pattern(N, E2, h);
    printf("N=%d edges %zu -> %zu bad=%d\n", N, E.size(), E2.size(), g_bad);
    return g_bad ? 1 : 0;
}
