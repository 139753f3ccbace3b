// Tie-exact top-K selection.
//
// NexDynIPC keeps the K nearest primitives per body pair with
//   std::nth_element(idx, idx + k, idx + n, score(a) < score(b)); idx.resize(k);
// then orders them with std::sort (ContactAssembler.cpp:67-88, idx = iota). When
// scores tie at the K boundary the kept SET depends on the standard library's
// introselect, and when kept scores tie their ORDER depends on its introsort. The
// oracle is built with GCC's libstdc++, so the device reproduces libstdc++'s published
// algorithms (Musser's introspective selection and sorting: median-of-three pivot moved
// to the front, unguarded Hoare partition, insertion sort on small ranges, heap
// selection / heap sort once the 2*floor(log2 n) depth budget is spent) step for step.
//
// The algorithm is written once over an element store: IdxStore permutes indices and
// compares their scores (the reference's formulation), KVStore / KV16Store move
// (score, index) pairs so that a comparison is a single load. All produce the same
// permutation. tk_nth_element_lr and tk_sort_blocks restate the two algorithms in the
// form the device runs them in parallel (one wave per list, geom.hip): partitions
// through the positions they swap, final insertion sorts as per-block rank sorts.
#pragma once
#include "dmath.h"

namespace nd {

struct IdxStore {
    int* a;
    const double* s;
    typedef int V;
    ND_HD V get(int i) const { return a[i]; }
    ND_HD void set(int i, V v) const { a[i] = v; }
    ND_HD bool less(V x, V y) const { return s[x] < s[y]; }
    ND_HD IdxStore offset(int o) const { return IdxStore{a + o, s}; }
};

struct KVPair {
    double k;
    int i;
};
struct KVStore {
    double* k;
    int* idx;
    typedef KVPair V;
    ND_HD V get(int i) const { KVPair p; p.k = k[i]; p.i = idx[i]; return p; }
    ND_HD void set(int i, V v) const { k[i] = v.k; idx[i] = v.i; }
    ND_HD bool less(V x, V y) const { return x.k < y.k; }
    ND_HD KVStore offset(int o) const { return KVStore{k + o, idx + o}; }
};

// (score, index) pairs with 16-bit indices (lists of at most 65536 primitives): the
// device's LDS-resident list.
struct KV16Store {
    double* k;
    unsigned short* idx;
    typedef KVPair V;
    ND_HD V get(int i) const { KVPair p; p.k = k[i]; p.i = idx[i]; return p; }
    ND_HD void set(int i, V v) const { k[i] = v.k; idx[i] = (unsigned short)v.i; }
    ND_HD bool less(V x, V y) const { return x.k < y.k; }
    ND_HD KV16Store offset(int o) const { return KV16Store{k + o, idx + o}; }
};

template <typename St>
ND_HD void tk_swap(const St& a, int i, int j) {
    const typename St::V t = a.get(i);
    a.set(i, a.get(j));
    a.set(j, t);
}

template <typename St>
ND_HD void tk_push_heap(const St& a, int first, int hole, int top, typename St::V value) {
    int parent = (hole - 1) / 2;
    while (hole > top && a.less(a.get(first + parent), value)) {
        a.set(first + hole, a.get(first + parent));
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a.set(first + hole, value);
}

template <typename St>
ND_HD void tk_adjust_heap(const St& a, int first, int hole, int len, typename St::V value) {
    const int top = hole;
    int second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (a.less(a.get(first + second), a.get(first + second - 1))) second--;
        a.set(first + hole, a.get(first + second));
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        a.set(first + hole, a.get(first + second - 1));
        hole = second - 1;
    }
    tk_push_heap(a, first, hole, top, value);
}

template <typename St>
ND_HD void tk_make_heap(const St& a, int first, int last) {
    const int len = last - first;
    if (len < 2) return;
    int parent = (len - 2) / 2;
    while (true) {
        tk_adjust_heap(a, first, parent, len, a.get(first + parent));
        if (parent == 0) return;
        parent--;
    }
}

template <typename St>
ND_HD void tk_heap_select(const St& a, int first, int middle, int last) {
    tk_make_heap(a, first, middle);
    for (int i = middle; i < last; ++i) {
        if (a.less(a.get(i), a.get(first))) {
            const typename St::V v = a.get(i);
            a.set(i, a.get(first));
            tk_adjust_heap(a, first, 0, middle - first, v);
        }
    }
}

template <typename St>
ND_HD void tk_move_median_to_first(const St& a, int result, int x, int y, int z) {
    if (a.less(a.get(x), a.get(y))) {
        if (a.less(a.get(y), a.get(z))) tk_swap(a, result, y);
        else if (a.less(a.get(x), a.get(z))) tk_swap(a, result, z);
        else tk_swap(a, result, x);
    } else if (a.less(a.get(x), a.get(z))) {
        tk_swap(a, result, x);
    } else if (a.less(a.get(y), a.get(z))) {
        tk_swap(a, result, z);
    } else {
        tk_swap(a, result, y);
    }
}

template <typename St>
ND_HD int tk_unguarded_partition(const St& a, int first, int last, int pivot) {
    const typename St::V pv = a.get(pivot);  // the pivot slot is not moved by this loop
    while (true) {
        while (a.less(a.get(first), pv)) ++first;
        --last;
        while (a.less(pv, a.get(last))) --last;
        if (!(first < last)) return first;
        tk_swap(a, first, last);
        ++first;
    }
}

template <typename St>
ND_HD void tk_insertion_sort(const St& a, int first, int last) {
    if (first == last) return;
    for (int i = first + 1; i != last; ++i) {
        const typename St::V v = a.get(i);
        if (a.less(v, a.get(first))) {
            for (int k = i; k > first; --k) a.set(k, a.get(k - 1));
            a.set(first, v);
        } else {
            int hole = i, next = i - 1;
            while (a.less(v, a.get(next))) {
                a.set(hole, a.get(next));
                hole = next;
                --next;
            }
            a.set(hole, v);
        }
    }
}

ND_HD int tk_floor_log2(int n) {
    int r = 0;
    while (n > 1) { n >>= 1; ++r; }
    return r;
}

template <typename St>
ND_HD void tk_unguarded_linear_insert(const St& a, int last) {
    const typename St::V v = a.get(last);
    int next = last - 1;
    while (a.less(v, a.get(next))) {
        a.set(last, a.get(next));
        last = next;
        --next;
    }
    a.set(last, v);
}

template <typename St>
ND_HD void tk_sort_heap(const St& a, int first, int last) {
    while (last - first > 1) {
        --last;
        const typename St::V v = a.get(last);
        a.set(last, a.get(first));
        tk_adjust_heap(a, first, 0, last - first, v);
    }
}

// sort(a, a + n) with the store's comparator: introsort as in libstdc++ (median-of-three
// quicksort down to ranges of 16, heap sort once the 2*floor(log2 n) depth budget is
// spent, then insertion sort over the first 16 and unguarded insertion of the rest).
// Partitioned ranges are independent, so an explicit stack replaces the recursion; the
// depths on it strictly decrease, so it never holds more than 2*log2(n) + 1 entries.
template <typename St>
ND_HD void tk_sort(const St& a, int n) {
    if (n < 2) return;
    int sf[64], sl[64], sd[64];
    int sp = 0;
    sf[0] = 0;
    sl[0] = n;
    sd[0] = 2 * tk_floor_log2(n);
    sp = 1;
    while (sp > 0) {
        --sp;
        const int first = sf[sp];
        int last = sl[sp], depth = sd[sp];
        while (last - first > 16) {
            if (depth == 0) {
                tk_heap_select(a, first, last, last);
                tk_sort_heap(a, first, last);
                break;
            }
            --depth;
            const int mid = first + (last - first) / 2;
            tk_move_median_to_first(a, first, first + 1, mid, last - 1);
            const int cut = tk_unguarded_partition(a, first + 1, last, first);
            sf[sp] = cut;
            sl[sp] = last;
            sd[sp] = depth;
            ++sp;
            last = cut;
        }
    }
    if (n > 16) {
        tk_insertion_sort(a, 0, 16);
        for (int i = 16; i < n; ++i) tk_unguarded_linear_insert(a, i);
    } else {
        tk_insertion_sort(a, 0, n);
    }
}

// The Hoare partition of tk_unguarded_partition(a, lo, hi, pivot) restated through the
// positions it swaps: L = ascending positions in [lo, hi) whose element is not less than
// the pivot, R = descending positions whose element is not greater (both taken before
// any swap). The k-th swap exchanges L_k and R_k for as long as L_k < R_k; the returned
// cut is min(L_{m+1}, R_m) for m swaps. Same permutation and cut as the scanning loop;
// the work-group version in geom.hip computes L and R with a parallel scan.
template <typename St>
ND_HD int tk_partition_lr(const St& a, int lo, int hi, int pivot, int* lpos, int* rpos) {
    const typename St::V pv = a.get(pivot);
    int nL = 0, nR = 0;
    for (int i = lo; i < hi; ++i)
        if (!a.less(a.get(i), pv)) lpos[nL++] = i;
    for (int i = hi - 1; i >= lo; --i)
        if (!a.less(pv, a.get(i))) rpos[nR++] = i;
    int m = 0;
    while (m < nL && m < nR && lpos[m] < rpos[m]) ++m;
    int cut = m < nL ? lpos[m] : hi;
    if (m >= 1 && rpos[m - 1] < cut) cut = rpos[m - 1];
    for (int q = 0; q < m; ++q) tk_swap(a, lpos[q], rpos[q]);
    return cut;
}

// Stable sort of every block of [0, n): blocks begin at the positions p with start[p]
// set (start[0] is set). Element i of block [s, e) moves to
//   s + #{j in [s, e): a_j < a_i} + #{j in [s, i): a_j equal to a_i}.
// libstdc++'s insertion sorts move an element left past strictly greater ones only, so
// they are stable sorts; the introsort loop leaves [0, n) as blocks in which every
// element is no greater than any element of a later block (Hoare partition around a
// pivot kept at the block's first position), so its final insertion sort over [0, n)
// moves no element out of its block and is this per-block stable sort. tmp holds n
// elements.
template <typename St>
ND_HD void tk_block_rank_sort(const St& a, int n, const unsigned char* start, typename St::V* tmp) {
    for (int i = 0; i < n; ++i) {
        int s = i, e = i + 1;
        while (!start[s]) --s;
        while (e < n && !start[e]) ++e;
        const typename St::V v = a.get(i);
        int r = s;
        for (int j = s; j < e; ++j) {
            const typename St::V u = a.get(j);
            r += a.less(u, v) || (j < i && !a.less(v, u));
        }
        tmp[r] = v;
    }
    for (int i = 0; i < n; ++i) a.set(i, tmp[i]);
}

// The first q in [0, mm) with lpos[q] >= rpos[q], or mm if there is none, for an
// ascending lpos and a descending rpos (the test is monotone in q), found as the device
// finds it with W = 64 lanes: probes at q = j * step (j < nq, step = ceil(mm / W),
// nq = ceil(mm / step)) bracket it to ((f - 1) * step, min(f * step, mm)], f the first
// probe that holds (nq if none), and rounds of W probes through the bracket (fewer than
// step positions; one round up to mm = 4096) find it.
template <typename T>
ND_HD int tk_first_crossing(const T* lpos, const T* rpos, int mm) {
    const int W = 64;
    const int step = (mm + W - 1) / W;
    const int nq = step ? (mm + step - 1) / step : 0;
    int f = nq;
    for (int j = 0; j < nq; ++j)
        if (lpos[j * step] >= rpos[j * step]) {
            f = j;
            break;
        }
    if (f == 0) return 0;
    const int qa = (f - 1) * step + 1, qb = f * step < mm ? f * step : mm;
    for (int base = qa; base < qb; base += W)
        for (int j = 0; j < W; ++j)
            if (base + j < qb && lpos[base + j] >= rpos[base + j]) return base + j;
    return qb;
}

// tk_nth_element with every partition done by tk_partition_lr and the final insertion
// sort of the last (at most 3-element) range done as a one-block rank sort.
template <typename St>
ND_HD void tk_nth_element_lr(const St& a, int nth, int n, int* lpos, int* rpos) {
    if (n == 0 || nth == n) return;
    int first = 0, last = n;
    int depth = tk_floor_log2(n) * 2;
    while (last - first > 3) {
        if (depth == 0) {
            tk_heap_select(a, first, nth + 1, last);
            tk_swap(a, first, nth);
            return;
        }
        --depth;
        const int mid = first + (last - first) / 2;
        tk_move_median_to_first(a, first, first + 1, mid, last - 1);
        const int cut = tk_partition_lr(a, first + 1, last, first, lpos, rpos);
        if (cut <= nth) first = cut;
        else last = cut;
    }
    typename St::V tmp[3];
    unsigned char start[3] = {1, 0, 0};
    const St sub = a.offset(first);
    tk_block_rank_sort(sub, last - first, start, tmp);
}

// tk_sort restated with the device's steps: partitions by tk_partition_lr, the final
// ranges of the introsort loop (at most 16 elements) marked as blocks, every element of
// a heap-sorted range a block of its own, and the final insertion sort done by
// tk_block_rank_sort. Same permutation as tk_sort. start and tmp hold n entries.
template <typename St>
ND_HD void tk_sort_blocks(const St& a, int n, int* lpos, int* rpos, unsigned char* start,
                          typename St::V* tmp) {
    if (n < 2) return;
    for (int i = 0; i < n; ++i) start[i] = 0;
    int sf[64], sl[64], sd[64];
    int sp = 1;
    sf[0] = 0;
    sl[0] = n;
    sd[0] = 2 * tk_floor_log2(n);
    while (sp > 0) {
        --sp;
        const int first = sf[sp];
        int last = sl[sp], depth = sd[sp];
        bool heap = false;
        while (last - first > 16) {
            if (depth == 0) {
                tk_heap_select(a, first, last, last);
                tk_sort_heap(a, first, last);
                for (int i = first; i < last; ++i) start[i] = 1;
                heap = true;
                break;
            }
            --depth;
            const int mid = first + (last - first) / 2;
            tk_move_median_to_first(a, first, first + 1, mid, last - 1);
            const int cut = tk_partition_lr(a, first + 1, last, first, lpos, rpos);
            sf[sp] = cut;
            sl[sp] = last;
            sd[sp] = depth;
            ++sp;
            last = cut;
        }
        if (!heap && first < n) start[first] = 1;
    }
    tk_block_rank_sort(a, n, start, tmp);
}

// nth_element(a, a + nth, a + n) with the store's comparator.
template <typename St>
ND_HD void tk_nth_element(const St& a, int nth, int n) {
    if (n == 0 || nth == n) return;
    int first = 0, last = n;
    int depth = tk_floor_log2(n) * 2;
    while (last - first > 3) {
        if (depth == 0) {
            tk_heap_select(a, first, nth + 1, last);
            tk_swap(a, first, nth);
            return;
        }
        --depth;
        const int mid = first + (last - first) / 2;
        tk_move_median_to_first(a, first, first + 1, mid, last - 1);
        const int cut = tk_unguarded_partition(a, first + 1, last, first);
        if (cut <= nth) first = cut;
        else last = cut;
    }
    tk_insertion_sort(a, first, last);
}

}  // namespace nd
