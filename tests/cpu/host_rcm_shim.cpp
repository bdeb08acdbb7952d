// Exposes the library's host RCM (fsd_amd/csrc/rcm.cpp) to the CPU tests.
void host_rcm(int N, const int *phead, const int *plist, int *perm);
extern "C" void test_host_rcm(int N, const int *phead, const int *plist, int *perm) { host_rcm(N, phead, plist, perm); }
