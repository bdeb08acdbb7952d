// CPU emulation of trsv_sweeps_kernel (fsd_amd/csrc/k_precond.hip): the
// schedules of sweep_plan.cpp, the packed operands as the pack kernels lay
// them out, and the kernel's thread mapping (16 waves x 21 rows of 3 lanes,
// dependency positions in batches of 8, values in batches of 4, 32-bit byte
// offsets) replayed level by level.  Checks every offset against its
// buffer, that no row reads a dependency before it is written, and the
// result against a direct block solve.
// usage: sweep_emulation N band max_lower_degree
#include "../../fsd_amd/csrc/sweep_plan.h"
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <algorithm>
#include <random>
#include <set>
int main(int argc, char **argv) {
    int N = argc > 1 ? atoi(argv[1]) : 3000; int band = argc > 2 ? atoi(argv[2]) : 40;
    std::mt19937 rng(5);
    // random symmetric graph, banded (like RCM output)
    std::vector<std::set<int>> adj(N);
    for (int P = 0; P < N; ++P) { int deg = rng() % (argc > 3 ? atoi(argv[3]) : 5); for (int t = 0; t < deg; ++t) { int Q = P - 1 - (int)(rng() % band); if (Q >= 0) { adj[P].insert(Q); adj[Q].insert(P);} } }
    std::vector<int> lrow(N + 1), lcol, urow(N + 1, 0), ucol, upos;
    for (int P = 0; P < N; ++P) { lrow[P] = lcol.size(); for (int Q : adj[P]) if (Q < P) { lcol.push_back(Q); urow[Q + 1]++; } }
    lrow[N] = lcol.size();
    for (int P = 0; P < N; ++P) urow[P + 1] += urow[P];
    ucol.resize(lcol.size()); upos.resize(lcol.size());
    { std::vector<int> fill(urow.begin(), urow.end() - 1);
      for (int P = 0; P < N; ++P) for (int e = lrow[P]; e < lrow[P + 1]; ++e) { int Q = lcol[e]; ucol[fill[Q]] = P; upos[fill[Q]] = e; fill[Q]++; } }
    int64_t nlow = lcol.size();
    std::vector<int> lf(N, 0), lb(N, 0); int mf = 0, mb = 0;
    for (int P = 0; P < N; ++P) { int l = 0; for (int e = lrow[P]; e < lrow[P + 1]; ++e) l = std::max(l, lf[lcol[e]] + 1); lf[P] = l; mf = std::max(mf, l); }
    for (int P = N - 1; P >= 0; --P) { int l = 0; for (int e = urow[P]; e < urow[P + 1]; ++e) l = std::max(l, lb[ucol[e]] + 1); lb[P] = l; mb = std::max(mb, l); }
    SweepPlan f, b;
    build_sweep_plan(N, lf.data(), mf + 1, lrow.data(), lcol.data(), nullptr, f);
    build_sweep_plan(N, lb.data(), mb + 1, urow.data(), ucol.data(), upos.data(), b);
    // random factor: lower blocks and lower-triangular diagonal blocks (well conditioned)
    std::uniform_real_distribution<double> U(-0.3, 0.3);
    std::vector<double> Lblk(36 * nlow), Ldiag(36 * N, 0.0);
    for (auto &v : Lblk) v = U(rng);
    for (int P = 0; P < N; ++P) for (int r = 0; r < 6; ++r) for (int c = 0; c <= r; ++c) Ldiag[36 * P + 6 * r + c] = (r == c) ? 1.5 + U(rng) : U(rng);
    std::vector<int> perm(N); for (int i = 0; i < N; ++i) perm[i] = i; std::shuffle(perm.begin(), perm.end(), rng);
    std::vector<double> bvec[2]; for (int r = 0; r < 2; ++r) { bvec[r].resize(6 * N); for (auto &v : bvec[r]) v = U(rng); }
    // direct: y = L^-1 Pb, x = L^-T y, out[perm[P]] = x[P]
    std::vector<double> ref[2];
    for (int r = 0; r < 2; ++r) {
        std::vector<double> y(6 * N), x(6 * N);
        for (int P = 0; P < N; ++P) {
            double v[6]; for (int a = 0; a < 6; ++a) v[a] = bvec[r][6 * perm[P] + a];
            for (int e = lrow[P]; e < lrow[P + 1]; ++e) for (int a = 0; a < 6; ++a) for (int c = 0; c < 6; ++c) v[a] -= Lblk[36 * e + 6 * a + c] * y[6 * lcol[e] + c];
            for (int a = 0; a < 6; ++a) { double s = v[a]; for (int c = 0; c < a; ++c) s -= Ldiag[36 * P + 6 * a + c] * y[6 * P + c]; y[6 * P + a] = s / Ldiag[36 * P + 7 * a]; }
        }
        for (int P = N - 1; P >= 0; --P) {
            double v[6]; for (int a = 0; a < 6; ++a) v[a] = y[6 * P + a];
            for (int u = urow[P]; u < urow[P + 1]; ++u) { int Q = ucol[u], e = upos[u]; for (int a = 0; a < 6; ++a) for (int c = 0; c < 6; ++c) v[a] -= Lblk[36 * e + 6 * c + a] * x[6 * Q + c]; }
            for (int a = 5; a >= 0; --a) { double s = v[a]; for (int c = a + 1; c < 6; ++c) s -= Ldiag[36 * P + 6 * c + a] * x[6 * P + c]; x[6 * P + a] = s / Ldiag[36 * P + 7 * a]; }
        }
        ref[r].resize(6 * N); for (int P = 0; P < N; ++P) for (int a = 0; a < 6; ++a) ref[r][6 * perm[P] + a] = x[6 * P + a];
    }
    // packed operands, as the pack kernels build them
    std::vector<double> fblk(36 * nlow), bblk(36 * nlow), fd(36 * N), bd(36 * N);
    for (int64_t e = 0; e < nlow; ++e) for (int q = 0; q < 36; ++q) { int a = q / 6, c = q % 6;
        if (f.esrc[e] < 0 || f.esrc[e] >= nlow || b.esrc[e] < 0 || b.esrc[e] >= nlow) { printf("esrc OOB\n"); return 1; }
        fblk[36 * e + q] = Lblk[36 * (int64_t)f.esrc[e] + q]; bblk[36 * e + q] = Lblk[36 * (int64_t)b.esrc[e] + 6 * c + a]; }
    for (int t = 0; t < 2 * N; ++t) { bool fw = t < N; int k = fw ? t : t - N; int P = fw ? f.order[k] : b.order[k];
        double Lm[36], Z[36] = {0}; for (int q = 0; q < 36; ++q) Lm[q] = Ldiag[36 * P + q];
        for (int c = 0; c < 6; ++c) { Z[7 * c] = 1.0 / Lm[7 * c]; for (int r = c + 1; r < 6; ++r) { double s = 0; for (int q = c; q < r; ++q) s += Lm[6 * r + q] * Z[6 * q + c]; Z[6 * r + c] = -s / Lm[7 * r]; } }
        double *dst = (fw ? fd.data() : bd.data()) + 36 * k; for (int a = 0; a < 6; ++a) for (int c = 0; c < 6; ++c) dst[6 * a + c] = fw ? Z[6 * a + c] : Z[6 * c + a]; }
    // table as uploaded
    std::vector<int> tab; int off[2][4];
    auto put = [&](const SweepPlan &pl, int *o) { o[0] = pl.nlev; o[1] = tab.size(); tab.insert(tab.end(), pl.levrow.begin(), pl.levrow.end()); o[2] = tab.size(); tab.insert(tab.end(), pl.levjd.begin(), pl.levjd.end()); o[3] = tab.size(); tab.insert(tab.end(), pl.rec.begin(), pl.rec.end()); };
    put(f, off[0]); put(b, off[1]);
    std::vector<int> fin(N), fyb(N), bout(N);
    for (int k = 0; k < N; ++k) { fin[k] = perm[f.order[k]]; fyb[k] = b.pos[f.order[k]]; bout[k] = perm[b.order[k]]; }
    const int64_t N6 = 6 * (int64_t)N;
    // gather
    std::vector<double> bin(12 * N), yf(12 * N, NAN), yb(12 * N, NAN), x(12 * N, NAN), out0(6 * N, NAN), out1(6 * N, NAN);
    for (int64_t t = 0; t < N6; ++t) { int64_t k = t / 6; int a = t % 6; bin[t] = bvec[0][6 * fin[k] + a]; bin[N6 + t] = bvec[1][6 * fin[k] + a]; }
    std::vector<char> wf(N, 0), wb(N, 0);
    long bad = 0, oob = 0;
    auto chk = [&](uint64_t byteoff, uint64_t bytes, uint64_t size) { if (byteoff + bytes > size) oob++; };
    const int JC = 8, JB = 4, nslot = 16 * 21;
    for (int sw = 0; sw < 2; ++sw) {
        int *o = off[sw]; const std::vector<int> &col = sw ? b.ecol : f.ecol; const std::vector<double> &blk = sw ? bblk : fblk, &dinv = sw ? bd : fd;
        const std::vector<double> &in = sw ? yb : bin; std::vector<double> &res = sw ? x : yf; std::vector<char> &wr = sw ? wb : wf;
        const std::vector<int> &scat = sw ? bout : fyb;
        for (int l = 0; l < o[0]; ++l) {
            int r0 = tab[o[1] + l], r1 = tab[o[1] + l + 1], j0 = tab[o[2] + l], nj = tab[o[2] + l + 1] - j0;
            std::vector<std::pair<int, std::array<double, 12>>> pend;
            for (int base = r0; base < r1; base += nslot) for (int slot = 0; slot < nslot; ++slot) {
                int k = base + slot; if (k >= r1) continue; int kl = k - r0;
                std::array<double, 12> w{};
                double v[2][6];
                for (int c = 0; c < 3; ++c) {
                    uint32_t kofs = 24u * k + 8u * c;
                    chk(kofs, 8, 4 * 2 * N6); chk(4 * N6 + kofs, 8, 4 * 2 * N6);
                    for (int i = 0; i < 2; ++i) { v[0][2 * c + i] = in[(kofs / 4) + i]; v[1][2 * c + i] = in[N6 + (kofs / 4) + i]; }
                    uint32_t dofs = 144u * k + 48u * c; chk(dofs, 48, 4 * 36 * (uint64_t)N);
                    for (int jb = 0; jb < nj; jb += JC) {
                        uint32_t eo[JC], qo[JC]; int n = 0;
                        for (int t = 0; t < JC; ++t) { int j = j0 + jb + t; bool has = (jb + t < nj) && kl < tab[o[3] + 2 * j + 1];
                            n += has; uint32_t e = has ? (uint32_t)(tab[o[3] + 2 * j] + kl) : 0u;
                            chk(4u * e, 4, 4 * (uint64_t)nlow);
                            eo[t] = 144u * e + 48u * c; qo[t] = 24u * (uint32_t)col[e]; }
                        for (int tb = 0; tb < JC; tb += JB) { if (tb >= n) break;
                            for (int t = tb; t < tb + JB; ++t) {
                                chk(eo[t], 48, 4 * 36 * (uint64_t)nlow); chk(qo[t], 24, 4 * 2 * N6); chk(4 * N6 + qo[t], 24, 4 * 2 * N6);
                                if (t >= n) continue;
                                int q = qo[t] / 24; if (!wr[q]) bad++;
                                int64_t e = eo[t] / 144;
                                for (int i = 0; i < 2; ++i) for (int m = 0; m < 6; ++m) { double L = blk[36 * e + 6 * (2 * c + i) + m];
                                    v[0][2 * c + i] -= L * res[6 * q + m]; v[1][2 * c + i] -= L * res[N6 + 6 * q + m]; }
                            } }
                    }
                }
                for (int a = 0; a < 6; ++a) for (int m = 0; m < 6; ++m) { w[a] += dinv[36 * (int64_t)k + 6 * a + m] * v[0][m]; w[6 + a] += dinv[36 * (int64_t)k + 6 * a + m] * v[1][m]; }
                chk(4u * k, 4, 4 * (uint64_t)N);
                pend.push_back({k, w});
            }
            for (auto &pw : pend) { int k = pw.first; for (int a = 0; a < 6; ++a) { res[6 * k + a] = pw.second[a]; res[N6 + 6 * k + a] = pw.second[6 + a];
                    double *c0 = sw ? out0.data() : yb.data(); double *c1 = sw ? out1.data() : yb.data() + N6;
                    c0[6 * scat[k] + a] = pw.second[a]; c1[6 * scat[k] + a] = pw.second[6 + a]; } wr[k] = 1; }
        }
    }
    if (oob) { printf("OUT OF BOUNDS accesses: %ld\n", oob); return 1; }
    double err = 0, mx = 0;
    for (int i = 0; i < 6 * N; ++i) { err = std::max(err, std::fabs(out0[i] - ref[0][i])); err = std::max(err, std::fabs(out1[i] - ref[1][i])); mx = std::max(mx, std::fabs(ref[0][i])); }
    printf("N=%d nlow=%ld levels f=%d b=%d tab=%zu ints, unready reads=%ld, oob=%ld, max err=%.3e (max |x|=%.3e)\n", N, (long)nlow, f.nlev, b.nlev, tab.size(), bad, oob, err, mx);
    return (bad || !(err < 1e-9 * std::max(1.0, mx))) ? 1 : 0;
}
