// Host check of the device top-K selection restatement (nexdynipc_amd/csrc/topk_select.h)
// against the C++ standard library's std::nth_element (libstdc++) on tie-heavy inputs.
// Prints "OK <cases>" or the first mismatch.
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>
#include "../../nexdynipc_amd/csrc/topk_select.h"

int main() {
    std::mt19937_64 rng(12345);
    long cases = 0;
    for (int trial = 0; trial < 40000; ++trial) {
        const int n = 1 + int(rng() % 2100);
        const int k = 1 + int(rng() % std::min(n, 40));
        const int levels = 1 + int(rng() % 6);
        std::vector<double> s(n);
        for (int i = 0; i < n; ++i) {
            const int mode = int(rng() % 4);
            if (mode == 0) s[i] = 0.0;
            else if (mode == 1) s[i] = double(rng() % levels) * 0.125;
            else s[i] = double(rng() % 1000000) * 1e-6;
        }
        if (trial % 7 == 0) std::sort(s.begin(), s.end());
        if (trial % 11 == 0) std::sort(s.rbegin(), s.rend());
        std::vector<int> a(n), b(n), c(n);
        std::vector<double> ck(s);
        std::iota(a.begin(), a.end(), 0);
        std::iota(b.begin(), b.end(), 0);
        std::iota(c.begin(), c.end(), 0);
        if (k < n) {
            std::nth_element(a.begin(), a.begin() + k, a.end(), [&](int x, int y) { return s[x] < s[y]; });
            nd::tk_nth_element(nd::IdxStore{b.data(), s.data()}, k, n);
            nd::tk_nth_element(nd::KVStore{ck.data(), c.data()}, k, n);
        }
        if (a != b || a != c) {
            std::printf("MISMATCH trial %d n %d k %d\n", trial, n, k);
            return 1;
        }
        ++cases;
    }
    std::printf("OK %ld\n", cases);
    return 0;
}
