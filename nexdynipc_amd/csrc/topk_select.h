// Tie-exact top-K selection.
//
// NexDynIPC keeps the K nearest primitives per body pair with
//   std::nth_element(idx, idx + k, idx + n, score(a) < score(b)); idx.resize(k);
// (ContactAssembler.cpp:67-88, idx = iota). When scores tie at the K boundary the kept
// SET depends on the standard library's introselect. The oracle is built with GCC's
// libstdc++, so the device reproduces libstdc++'s published introselect (Musser's
// introspective selection: median-of-three pivot moved to the front, unguarded Hoare
// partition, insertion sort below 4 elements, heap selection once the 2*floor(log2 n)
// depth budget is spent) step for step on the index array. It runs on one lane over
// an LDS-resident list and is only entered when a tie straddles the boundary; the
// common case is decided by a parallel sort (see candidates.hip).
#pragma once
#include "dmath.h"

namespace nd {

struct TopkCmp {
    const double* s;
    ND_HD bool operator()(int a, int b) const { return s[a] < s[b]; }
};

ND_HD void tk_swap(int* a, int i, int j) {
    const int t = a[i];
    a[i] = a[j];
    a[j] = t;
}

ND_HD void tk_push_heap(int* a, int first, int hole, int top, int value, const TopkCmp& c) {
    int parent = (hole - 1) / 2;
    while (hole > top && c(a[first + parent], value)) {
        a[first + hole] = a[first + parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[first + hole] = value;
}

ND_HD void tk_adjust_heap(int* a, int first, int hole, int len, int value, const TopkCmp& c) {
    const int top = hole;
    int second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (c(a[first + second], a[first + second - 1])) second--;
        a[first + hole] = a[first + second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        a[first + hole] = a[first + second - 1];
        hole = second - 1;
    }
    tk_push_heap(a, first, hole, top, value, c);
}

ND_HD void tk_make_heap(int* a, int first, int last, const TopkCmp& c) {
    const int len = last - first;
    if (len < 2) return;
    int parent = (len - 2) / 2;
    while (true) {
        const int v = a[first + parent];
        tk_adjust_heap(a, first, parent, len, v, c);
        if (parent == 0) return;
        parent--;
    }
}

ND_HD void tk_heap_select(int* a, int first, int middle, int last, const TopkCmp& c) {
    tk_make_heap(a, first, middle, c);
    for (int i = middle; i < last; ++i) {
        if (c(a[i], a[first])) {
            const int v = a[i];
            a[i] = a[first];
            tk_adjust_heap(a, first, 0, middle - first, v, c);
        }
    }
}

ND_HD void tk_move_median_to_first(int* a, int result, int x, int y, int z, const TopkCmp& c) {
    if (c(a[x], a[y])) {
        if (c(a[y], a[z])) tk_swap(a, result, y);
        else if (c(a[x], a[z])) tk_swap(a, result, z);
        else tk_swap(a, result, x);
    } else if (c(a[x], a[z])) {
        tk_swap(a, result, x);
    } else if (c(a[y], a[z])) {
        tk_swap(a, result, z);
    } else {
        tk_swap(a, result, y);
    }
}

ND_HD int tk_unguarded_partition(int* a, int first, int last, int pivot, const TopkCmp& c) {
    while (true) {
        while (c(a[first], a[pivot])) ++first;
        --last;
        while (c(a[pivot], a[last])) --last;
        if (!(first < last)) return first;
        tk_swap(a, first, last);
        ++first;
    }
}

ND_HD void tk_insertion_sort(int* a, int first, int last, const TopkCmp& c) {
    if (first == last) return;
    for (int i = first + 1; i != last; ++i) {
        const int v = a[i];
        if (c(v, a[first])) {
            for (int k = i; k > first; --k) a[k] = a[k - 1];
            a[first] = v;
        } else {
            int hole = i, next = i - 1;
            while (c(v, a[next])) {
                a[hole] = a[next];
                hole = next;
                --next;
            }
            a[hole] = v;
        }
    }
}

ND_HD int tk_floor_log2(int n) {
    int r = 0;
    while (n > 1) { n >>= 1; ++r; }
    return r;
}

// nth_element(a, a + nth, a + n) with comparator score[x] < score[y].
ND_HD void tk_nth_element(int* a, int nth, int n, const TopkCmp& c) {
    if (n == 0 || nth == n) return;
    int first = 0, last = n;
    int depth = tk_floor_log2(n) * 2;
    while (last - first > 3) {
        if (depth == 0) {
            tk_heap_select(a, first, nth + 1, last, c);
            tk_swap(a, first, nth);
            return;
        }
        --depth;
        const int mid = first + (last - first) / 2;
        tk_move_median_to_first(a, first, first + 1, mid, last - 1, c);
        const int cut = tk_unguarded_partition(a, first + 1, last, first, c);
        if (cut <= nth) first = cut;
        else last = cut;
    }
    tk_insertion_sort(a, first, last, c);
}

}  // namespace nd
