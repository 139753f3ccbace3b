// Host check of the device top-K selection restatement (nexdynipc_amd/csrc/topk_select.h)
// against the C++ standard library (libstdc++) on tie-heavy inputs: std::nth_element
// followed by std::sort of the kept prefix (the reference's select_top_k_indices), and
// std::sort of whole arrays. The complete permutations must agree, not only the sets.
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
        std::vector<int> a(n), b(n), c(n), e(n), lp(n), rp(n);
        std::vector<double> ck(s), ek(s);
        std::iota(a.begin(), a.end(), 0);
        std::iota(b.begin(), b.end(), 0);
        std::iota(c.begin(), c.end(), 0);
        std::iota(e.begin(), e.end(), 0);
        auto cmp = [&](int x, int y) { return s[x] < s[y]; };
        if (k < n) {
            std::nth_element(a.begin(), a.begin() + k, a.end(), cmp);
            nd::tk_nth_element(nd::IdxStore{b.data(), s.data()}, k, n);
            nd::tk_nth_element(nd::KVStore{ck.data(), c.data()}, k, n);
            nd::tk_nth_element_lr(nd::KVStore{ek.data(), e.data()}, k, n, lp.data(), rp.data());
        }
        if (a != e) {
            std::printf("MISMATCH partition-by-positions trial %d n %d k %d\n", trial, n, k);
            return 1;
        }
        if (a != b || a != c) {
            std::printf("MISMATCH nth_element trial %d n %d k %d\n", trial, n, k);
            return 1;
        }
        std::sort(a.begin(), a.begin() + k, cmp);
        nd::tk_sort(nd::IdxStore{b.data(), s.data()}, k);
        nd::tk_sort(nd::KVStore{ck.data(), c.data()}, k);
        if (a != b || a != c) {
            std::printf("MISMATCH top-k order trial %d n %d k %d\n", trial, n, k);
            return 1;
        }
        ++cases;
    }
    // whole-array sorts: random, presorted, reversed, constant, few levels
    for (int trial = 0; trial < 4000; ++trial) {
        const int n = 1 + int(rng() % 3000);
        const int levels = 1 + int(rng() % 8);
        std::vector<double> s(n);
        for (int i = 0; i < n; ++i) s[i] = double(rng() % levels);
        if (trial % 5 == 1) std::sort(s.begin(), s.end());
        if (trial % 5 == 2) std::sort(s.rbegin(), s.rend());
        if (trial % 5 == 3) for (int i = 0; i < n; ++i) s[i] = double((i * 7919) % n);
        std::vector<int> a(n), b(n), c(n);
        std::vector<double> ck(s);
        std::iota(a.begin(), a.end(), 0);
        std::iota(b.begin(), b.end(), 0);
        std::iota(c.begin(), c.end(), 0);
        std::sort(a.begin(), a.end(), [&](int x, int y) { return s[x] < s[y]; });
        nd::tk_sort(nd::IdxStore{b.data(), s.data()}, n);
        nd::tk_sort(nd::KVStore{ck.data(), c.data()}, n);
        if (a != b || a != c) {
            std::printf("MISMATCH sort trial %d n %d\n", trial, n);
            return 1;
        }
        ++cases;
    }
    std::printf("OK %ld\n", cases);
    return 0;
}
