// Host check of the device top-K selection restatement (nexdynipc_amd/csrc/topk_select.h)
// against the C++ standard library (libstdc++) on tie-heavy inputs: std::nth_element
// followed by std::sort of the kept prefix (the reference's select_top_k_indices), and
// std::sort of whole arrays. The complete permutations must agree, not only the sets.
// The device's formulation (tk_nth_element_lr, tk_sort_blocks: partitions through the
// positions they swap, final insertion sorts as per-block rank sorts) is checked too.
// Prints "OK <cases>" or the first mismatch.
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <string>
#include <vector>
#include "../../nexdynipc_amd/csrc/topk_select.h"

// McIlroy's adversary ("A killer adversary for quicksort", 1999): values are fixed
// lazily as the standard library compares them, so that every pivot it picks is among
// the smallest remaining values. The resulting array drives introsort and introselect
// into their depth limits (the heap sort / heap select branches).
static std::vector<double> killer_input(int n, int k) {
    std::vector<int> val(n, n), ptr(n);
    std::iota(ptr.begin(), ptr.end(), 0);
    int nsolid = 0, candidate = 0;
    auto cmp = [&](int x, int y) {
        if (val[x] == n && val[y] == n) {
            if (x == candidate) val[x] = nsolid++;
            else val[y] = nsolid++;
        }
        if (val[x] == n) candidate = x;
        else if (val[y] == n) candidate = y;
        return val[x] < val[y];
    };
    if (k < n) std::nth_element(ptr.begin(), ptr.begin() + k, ptr.end(), cmp);
    else std::sort(ptr.begin(), ptr.end(), cmp);
    std::vector<double> s(n);
    for (int i = 0; i < n; ++i) s[i] = double(val[i]);
    return s;
}

static bool check_case(const std::vector<double>& s, int k) {
    const int n = int(s.size());
    std::vector<int> a(n), e(n), lp(n), rp(n);
    std::vector<double> ek(s);
    std::vector<unsigned char> st(n);
    std::vector<nd::KVPair> tmp(n);
    std::iota(a.begin(), a.end(), 0);
    std::iota(e.begin(), e.end(), 0);
    auto cmp = [&](int x, int y) { return s[x] < s[y]; };
    if (k < n) {
        std::nth_element(a.begin(), a.begin() + k, a.end(), cmp);
        nd::tk_nth_element_lr(nd::KVStore{ek.data(), e.data()}, k, n, lp.data(), rp.data());
    }
    std::sort(a.begin(), a.begin() + k, cmp);
    nd::tk_sort_blocks(nd::KVStore{ek.data(), e.data()}, k, lp.data(), rp.data(), st.data(), tmp.data());
    return a == e;
}

int main(int argc, char** argv) {
    if (argc == 3 && std::string(argv[1]) == "--dump") {
        // the adversarial inputs as text: "n k" then n values, one case per line
        FILE* f = std::fopen(argv[2], "w");
        if (!f) return 2;
        for (int n : {17, 33, 64, 100, 257, 1000, 2048, 4096, 16384}) {
            for (int k : {n, 1, 5, 32, n / 2}) {
                if (k < 1 || k > n) continue;
                // beyond 2048 entries: the device path's kept counts (k <= 2048), few cases
                if (n > 2048 && !(k == 32 || (n == 4096 && (k == 1 || k == n / 2)))) continue;
                for (int target : {k, n}) {
                    const std::vector<double> s = killer_input(n, target);
                    std::fprintf(f, "%d %d", n, k);
                    for (double v : s) std::fprintf(f, " %d", int(v));
                    std::fprintf(f, "\n");
                }
            }
        }
        std::fclose(f);
        return 0;
    }
    std::mt19937_64 rng(12345);
    long cases = 0;
    // the device's search for the first crossing against a linear scan, on every
    // crossing position of every length up to the LDS list limit, then on sampled
    // lengths and positions up to the 65535-entry limit (several second-round passes)
    std::vector<int> mms;
    for (int mm = 0; mm <= 2048; ++mm) mms.push_back(mm);
    for (int mm = 2049; mm <= 65535; mm += 997) mms.push_back(mm);
    for (int mm : {4095, 4096, 4097, 8191, 8192, 65534, 65535}) mms.push_back(mm);
    for (int mm : mms) {
        for (int m = 0; m <= mm; m += (mm > 2048 ? 1 + mm / 400 : (mm > 300 ? 7 : 1))) {
            std::vector<int> lp(mm), rp(mm);
            for (int q = 0; q < mm; ++q) {
                lp[q] = 2 * q;
                rp[q] = q < m ? 2 * q + 1 + 2 * (mm - q) : 2 * q - 1;  // crosses at q = m
            }
            int lin = 0;
            while (lin < mm && lp[lin] < rp[lin]) ++lin;
            if (lin != m || nd::tk_first_crossing(lp.data(), rp.data(), mm) != m) {
                std::printf("MISMATCH first crossing mm %d m %d\n", mm, m);
                return 1;
            }
            ++cases;
        }
    }
    for (int n : {17, 33, 64, 100, 257, 1000, 2048, 4096, 16384}) {
        for (int k : {n, 1, 5, 32, n / 2}) {
            if (k < 1 || k > n) continue;
            if (!check_case(killer_input(n, k), k) || !check_case(killer_input(n, n), k)) {
                std::printf("MISMATCH adversarial input n %d k %d\n", n, k);
                return 1;
            }
            cases += 2;
        }
    }
    for (int trial = 0; trial < 40000; ++trial) {
        // one trial in 100 at a list length of the HBM path (2049 to 65535 entries)
        const int n = trial % 100 == 99 ? 2049 + int(rng() % 63487) : 1 + int(rng() % 2100);
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
        {
            std::vector<unsigned char> st(n);
            std::vector<nd::KVPair> tmp(n);
            nd::tk_sort_blocks(nd::KVStore{ek.data(), e.data()}, k, lp.data(), rp.data(), st.data(), tmp.data());
        }
        if (a != b || a != c || a != e) {
            std::printf("MISMATCH top-k order trial %d n %d k %d\n", trial, n, k);
            return 1;
        }
        ++cases;
    }
    // whole-array sorts: random, presorted, reversed, constant, few levels; one trial in
    // 50 at a length of the contact cap's sort (2001 to 40000 raw contacts), and the
    // adversary's input for std::sort at 40000
    for (int k : {40000, 2000}) {
        if (!check_case(killer_input(40000, 40000), k)) {
            std::printf("MISMATCH adversarial sort n 40000 k %d\n", k);
            return 1;
        }
        ++cases;
    }
    for (int trial = 0; trial < 4000; ++trial) {
        const int n = trial % 50 == 49 ? 2001 + int(rng() % 38000) : 1 + int(rng() % 3000);
        const int levels = 1 + int(rng() % 8);
        std::vector<double> s(n);
        for (int i = 0; i < n; ++i) s[i] = double(rng() % levels);
        if (trial % 5 == 1) std::sort(s.begin(), s.end());
        if (trial % 5 == 2) std::sort(s.rbegin(), s.rend());
        if (trial % 5 == 3) for (int i = 0; i < n; ++i) s[i] = double((i * 7919) % n);
        std::vector<int> a(n), b(n), c(n), e(n), lp(n), rp(n);
        std::vector<double> ck(s), ek(s);
        std::vector<unsigned char> st(n);
        std::vector<nd::KVPair> tmp(n);
        std::iota(a.begin(), a.end(), 0);
        std::iota(b.begin(), b.end(), 0);
        std::iota(c.begin(), c.end(), 0);
        std::iota(e.begin(), e.end(), 0);
        std::sort(a.begin(), a.end(), [&](int x, int y) { return s[x] < s[y]; });
        nd::tk_sort(nd::IdxStore{b.data(), s.data()}, n);
        nd::tk_sort(nd::KVStore{ck.data(), c.data()}, n);
        nd::tk_sort_blocks(nd::KVStore{ek.data(), e.data()}, n, lp.data(), rp.data(), st.data(), tmp.data());
        if (a != b || a != c || a != e) {
            std::printf("MISMATCH sort trial %d n %d\n", trial, n);
            return 1;
        }
        ++cases;
    }
    std::printf("OK %ld\n", cases);
    return 0;
}
