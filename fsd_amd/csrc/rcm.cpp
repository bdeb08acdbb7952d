// Reverse Cuthill-McKee ordering of the pruned particle graph, following
// George & Liu's SPARSPAK GENRCM (the routine the reference calls at
// Precondition.cu:1009; rcm.cpp:870-1120, 2211-2262, 2340-2477, 3152-3574).
// The traversal order and tie-breaking are kept so the permutation is the
// reference's: degrees counted with the self loop the pruned list carries,
// pseudo-peripheral root from the last BFS level, and the children of a node
// ordered as rcm.cpp:3340-3362 orders them -- its insertion sort never
// compares the first new child (loop `while ( fnbr < l )`), so that child
// stays first and only the remaining children are sorted (stably) by
// increasing degree.  Indices are 0-based here.
//
// The graph is read in a fixed-stride layout (node v's neighbours at
// rows[v * stride ...], cnt[v] of them), the layout the pruned lists arrive
// in, so a node's row is one or two cache lines; the breadth-first loops
// prefetch the rows of queued nodes a few places ahead.
#include <algorithm>
#include <vector>

#include "precond_plan.h"

namespace {

constexpr int kAhead = 8;   // queue places prefetched ahead

struct Graph {
    const int *cnt;
    const int *rows;
    int stride;
    const int *row(int v) const { return rows + (size_t)v * stride; }
    void prefetch(int v) const
    {
        __builtin_prefetch(row(v));
        __builtin_prefetch(row(v) + stride - 1);
        __builtin_prefetch(cnt + v);
    }
};

// BFS from root over unmasked nodes into order[0, n) (order holds the
// component plus one slot); lstart gets the level starts.  Masks are restored
// on return; returns the number of levels.
int bfs_levels(const Graph &g, int root, char *mask, int *order, int &n, std::vector<int> &lstart)
{
    lstart.clear();
    order[0] = root;
    n = 1;
    mask[root] = 0;
    int head = 0;
    for (;;) {
        lstart.push_back(head);
        const int level_end = n;
        for (; head < level_end; ++head) {
            if (head + kAhead < n) g.prefetch(order[head + kAhead]);
            const int v = order[head];
            const int *r = g.row(v);
            const int c = g.cnt[v];
            for (int e = 0; e < c; ++e) {   // branch-free append of the unmasked neighbours
                const int w = r[e];
                const int m = mask[w];
                order[n] = w;
                mask[w] = 0;
                n += m;
            }
        }
        if (n == level_end) break;
    }
    lstart.push_back(n);
    for (int q = 0; q < n; ++q) mask[order[q]] = 1;
    return (int)lstart.size() - 1;
}

// FNROOT: a pseudo-peripheral node of root's component.  The degree of a node
// of the last level counts its unmasked neighbours; the whole component is
// unmasked while it is searched and has no neighbour outside it, so that is
// its full degree (cnt).
int pseudo_peripheral(const Graph &g, int root, char *mask, int *order, std::vector<int> &lstart)
{
    int size = 0;
    int nlev = bfs_levels(g, root, mask, order, size, lstart);
    if (nlev == 1 || nlev == size) return root;
    for (;;) {
        const int first_last = lstart[nlev - 1];
        int best = order[first_last];
        if (first_last < size - 1) {
            int mindeg = size;
            for (int q = first_last; q < size; ++q) {
                const int v = order[q];
                const int d = g.cnt[v];
                if (d < mindeg) { best = v; mindeg = d; }
            }
        }
        root = best;
        int size2 = 0;
        const int nlev2 = bfs_levels(g, root, mask, order, size2, lstart);
        if (nlev2 <= nlev) break;
        nlev = nlev2;
        if (size <= nlev) break;
    }
    return root;
}

}  // namespace

void host_rcm_rows(int N, const int *cnt, const int *rows, int stride, int *perm)
{
    Graph g{cnt, rows, stride};
    std::vector<char> mask(N, 1);
    std::vector<int> order(N + 1), lstart;
    int placed = 0;
    for (int s = 0; s < N && placed < N; ++s) {
        if (!mask[s]) continue;
        const int root = pseudo_peripheral(g, s, mask.data(), order.data(), lstart);
        // Cuthill-McKee into order: children of each node appended in increasing degree (the full
        // degree, as in the root search: the component is unnumbered while it is ordered)
        int *out = order.data();
        int n_out = 1;
        out[0] = root;
        mask[root] = 0;
        for (int q = 0; q < n_out; ++q) {
            if (q + kAhead < n_out) g.prefetch(out[q + kAhead]);
            const int v = out[q];
            const int first = n_out;
            const int *r = g.row(v);
            const int c = g.cnt[v];
            for (int e = 0; e < c; ++e) {
                const int w = r[e];
                const int m = mask[w];
                out[n_out] = w;
                mask[w] = 0;
                n_out += m;
            }
            // stable insertion sort by increasing degree of all children but the first (a few at most)
            for (int k = first + 2; k < n_out; ++k) {
                const int v2 = out[k], d2 = cnt[v2];
                int m = k;
                while (m > first + 1 && cnt[out[m - 1]] > d2) { out[m] = out[m - 1]; --m; }
                out[m] = v2;
            }
        }
        std::reverse_copy(out, out + n_out, perm + placed);
        placed += n_out;
    }
}

void host_rcm(int N, const int *phead, const int *plist, int *perm)
{
    // packed lists to the fixed-stride layout
    int stride = 1;
    for (int v = 0; v < N; ++v) stride = std::max(stride, phead[v + 1] - phead[v]);
    std::vector<int> cnt(N), rows((size_t)N * stride);
    for (int v = 0; v < N; ++v) {
        cnt[v] = phead[v + 1] - phead[v];
        std::copy(plist + phead[v], plist + phead[v + 1], rows.begin() + (size_t)v * stride);
    }
    host_rcm_rows(N, cnt.data(), rows.data(), stride, perm);
}
