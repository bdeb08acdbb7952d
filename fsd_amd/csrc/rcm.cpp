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
#include <vector>
#include <algorithm>

namespace {

struct Graph {
    const int *xadj;
    const int *adj;
};

// BFS from root over unmasked nodes; fills order and level starts.  Masks are
// restored on return.
int bfs_levels(const Graph &g, int root, std::vector<char> &mask, std::vector<int> &order, std::vector<int> &lstart)
{
    order.clear();
    lstart.clear();
    order.push_back(root);
    mask[root] = 0;
    size_t head = 0;
    for (;;) {
        lstart.push_back((int)head);
        const size_t level_end = order.size();
        for (; head < level_end; ++head) {
            const int v = order[head];
            for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
                const int w = g.adj[e];
                if (mask[w]) { mask[w] = 0; order.push_back(w); }
            }
        }
        if (order.size() == level_end) break;
    }
    lstart.push_back((int)order.size());
    for (int v : order) mask[v] = 1;
    return (int)lstart.size() - 1;
}

int pseudo_peripheral(const Graph &g, int root, std::vector<char> &mask, std::vector<int> &order,
                      std::vector<int> &lstart)
{
    int nlev = bfs_levels(g, root, mask, order, lstart);
    const int size = (int)order.size();
    if (nlev == 1 || nlev == size) return root;
    for (;;) {
        const int first_last = lstart[nlev - 1];
        int best = order[first_last];
        if (first_last < size - 1) {
            int mindeg = size;
            for (int q = first_last; q < size; ++q) {
                const int v = order[q];
                int d = 0;
                for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) d += mask[g.adj[e]] ? 1 : 0;
                if (d < mindeg) { best = v; mindeg = d; }
            }
        }
        root = best;
        const int nlev2 = bfs_levels(g, root, mask, order, lstart);
        if (nlev2 <= nlev) break;
        nlev = nlev2;
        if (size <= nlev) break;
    }
    return root;
}

}  // namespace

void host_rcm(int N, const int *phead, const int *plist, int *perm)
{
    Graph g{phead, plist};
    std::vector<char> mask(N, 1);
    std::vector<int> deg(N, 0), order, lstart;
    for (int v = 0; v < N; ++v) deg[v] = phead[v + 1] - phead[v];
    order.reserve(N);
    int placed = 0;
    for (int s = 0; s < N && placed < N; ++s) {
        if (!mask[s]) continue;
        const int root = pseudo_peripheral(g, s, mask, order, lstart);
        // Cuthill-McKee: children of each node appended in increasing degree.  The degree counts
        // unmasked neighbours (self loop included); a component is unnumbered while it is ordered and
        // has no neighbour outside it, so that is the full degree of the packed list (deg above).
        int *out = perm + placed;
        int n_out = 1;
        out[0] = root;
        mask[root] = 0;
        for (int q = 0; q < n_out; ++q) {
            const int v = out[q];
            const int first = n_out;
            for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
                const int w = g.adj[e];
                if (mask[w]) { mask[w] = 0; out[n_out++] = w; }
            }
            // stable insertion sort by increasing degree of all children but the first (a few at most)
            for (int k = first + 2; k < n_out; ++k) {
                const int v2 = out[k], d2 = deg[v2];
                int m = k;
                while (m > first + 1 && deg[out[m - 1]] > d2) { out[m] = out[m - 1]; --m; }
                out[m] = v2;
            }
        }
        std::reverse(out, out + n_out);
        placed += n_out;
    }
}
