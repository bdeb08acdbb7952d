// C-ABI entry to the reference's RCM (source/rcm.cpp, compiled where it lies
// by oracle/build_ref.sh) with the call-site conventions of
// Precondition_Reorder (Precondition.cu:993-1012): 0-based packed adjacency
// in, 1-based for genrcm, 0-based permutation out.  Test infrastructure only.
#include <vector>

void genrcm(int node_num, int adj_num, int adj_row[], int adj[], int perm[]);

extern "C" int ref_genrcm(int node_num, int adj_num, const int *head0, const int *list0, int *perm0)
{
    std::vector<int> head(node_num + 1), list(adj_num > 0 ? adj_num : 1);
    for (int i = 0; i <= node_num; ++i) head[i] = head0[i] + 1;
    for (int i = 0; i < adj_num; ++i) list[i] = list0[i] + 1;
    genrcm(node_num, adj_num, head.data(), list.data(), perm0);
    for (int i = 0; i < node_num; ++i) perm0[i] -= 1;
    return 0;
}
