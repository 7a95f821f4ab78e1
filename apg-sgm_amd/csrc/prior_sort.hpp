// The order std::sort (libstdc++, GCC) leaves a range of (key, weight)
// pairs in when it compares keys only.  The prior's per-cell statistics sum
// weights in sorted order (reference src/prior_estimator.cpp:212-246, two
// std::sort calls with a key-only comparator), so elements with equal keys
// must end up exactly where libstdc++ puts them for the float sums to round
// alike.  The algorithm is libstdc++'s introsort (D. R. Musser, "Introspective
// Sorting and Selection Algorithms", Software: Practice and Experience 27(8),
// 1997), with that library's choices:
//   * while a range holds more than 16 elements: the median of the elements
//     at first + 1, the middle (first + n / 2) and last - 1 is swapped to
//     first, then a Hoare partition of [first + 1, last) around it (both scans
//     stop on keys equal to the pivot); the upper part is sorted first (order
//     does not matter: the parts are disjoint), the lower part continues;
//   * a range still longer than 16 once 2 * floor(log2 n) partitions deep is
//     heap-sorted (make_heap, then pop_heap down to one element, each sift
//     moving the hole to a leaf along the larger child and back up);
//   * a final insertion sort: plain over the first 16 elements, then
//     unguarded insertions (the partitions leave a minimum in front).
// Shared by the device kernels (k_prior.hip) and a host test that checks it
// against std::sort itself.
#pragma once

#if defined(__HIPCC__)
#define APG_HD __host__ __device__
#else
#define APG_HD
#endif

namespace apgk {

struct KeyW {
    float k, w;
};

namespace lsort {

APG_HD inline bool less(const KeyW& a, const KeyW& b) { return a.k < b.k; }
APG_HD inline void swap(KeyW* a, int i, int j) {
    const KeyW t = a[i];
    a[i] = a[j];
    a[j] = t;
}

// insertion of a[i] into a[0, i) that knows a smaller element lies in front
APG_HD inline void unguarded_insert(KeyW* a, int i) {
    const KeyW v = a[i];
    int j = i;
    while (less(v, a[j - 1])) {
        a[j] = a[j - 1];
        --j;
    }
    a[j] = v;
}

APG_HD inline void insertion(KeyW* a, int n) {
    for (int i = 1; i < n; ++i) {
        if (less(a[i], a[0])) {
            const KeyW v = a[i];
            for (int j = i; j > 0; --j) a[j] = a[j - 1];
            a[0] = v;
        } else {
            unguarded_insert(a, i);
        }
    }
}

// sift of value v from the hole down to a leaf along the larger child, then up
APG_HD inline void adjust_heap(KeyW* a, int hole, int len, KeyW v) {
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (less(a[child], a[child - 1])) --child;
        a[hole] = a[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        a[hole] = a[child - 1];
        hole = child - 1;
    }
    int parent = (hole - 1) / 2;
    while (hole > top && less(a[parent], v)) {
        a[hole] = a[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[hole] = v;
}

APG_HD inline void heap_sort(KeyW* a, int n) {
    if (n >= 2) {
        for (int parent = (n - 2) / 2;; --parent) {
            adjust_heap(a, parent, n, a[parent]);
            if (parent == 0) break;
        }
    }
    for (int len = n; len > 1;) {
        --len;
        const KeyW v = a[len];
        a[len] = a[0];
        adjust_heap(a, 0, len, v);
    }
}

// median of a[i], a[j], a[k] swapped to a[r]
APG_HD inline void median_to_first(KeyW* a, int r, int i, int j, int k) {
    if (less(a[i], a[j])) {
        if (less(a[j], a[k])) swap(a, r, j);
        else if (less(a[i], a[k])) swap(a, r, k);
        else swap(a, r, i);
    } else if (less(a[i], a[k])) {
        swap(a, r, i);
    } else if (less(a[j], a[k])) {
        swap(a, r, k);
    } else {
        swap(a, r, j);
    }
}

// Hoare partition of [lo, hi) around a[p]; returns the cut
APG_HD inline int partition(KeyW* a, int lo, int hi, int p) {
    for (;;) {
        while (less(a[lo], a[p])) ++lo;
        --hi;
        while (less(a[p], a[hi])) --hi;
        if (!(lo < hi)) return lo;
        swap(a, lo, hi);
        ++lo;
    }
}

APG_HD inline int floor_log2(int n) {
    int r = 0;
    while (n > 1) {
        n >>= 1;
        ++r;
    }
    return r;
}

}  // namespace lsort

// Sorts a[0, n) by key into the order std::sort leaves.
APG_HD inline void sort_like_libstdcxx(KeyW* a, int n) {
    if (n <= 1) return;
    constexpr int kRun = 16;
    // pending upper parts (lo, hi, depth): at most one per partition level
    int st_lo[64], st_hi[64], st_depth[64];
    int sp = 0;
    int lo = 0, hi = n, depth = 2 * lsort::floor_log2(n);
    for (;;) {
        while (hi - lo > kRun) {
            if (depth == 0) {
                lsort::heap_sort(a + lo, hi - lo);
                break;
            }
            --depth;
            lsort::median_to_first(a, lo, lo + 1, lo + (hi - lo) / 2, hi - 1);
            const int cut = lsort::partition(a, lo + 1, hi, lo);
            st_lo[sp] = cut;
            st_hi[sp] = hi;
            st_depth[sp] = depth;
            ++sp;
            hi = cut;
        }
        if (sp == 0) break;
        --sp;
        lo = st_lo[sp];
        hi = st_hi[sp];
        depth = st_depth[sp];
    }
    if (n > kRun) {
        lsort::insertion(a, kRun);
        for (int i = kRun; i < n; ++i) lsort::unguarded_insert(a, i);
    } else {
        lsort::insertion(a, n);
    }
}

}  // namespace apgk
