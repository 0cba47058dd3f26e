// nli_introsort.h -- the element order of libstdc++'s std::sort (GCC 11) on (key, value)
// pairs compared by key alone, for the weighted median of patch_non_local_medians.cpp:110-126.
//
// The reference sorts (colour, weight) pairs with std::sort and CompareByColor (key <
// key), which is not stable: pairs with equal colours end in an order that depends on the
// algorithm, and the weighted median subtracts the weights from a float total in that
// order, so the rounding of the running total, and in rare cases the median, depend on it.
// This is a restatement of the published algorithm of libstdc++'s std::sort (introsort,
// Musser 1997, as bits/stl_algo.h and bits/stl_heap.h implement it): quicksort with the
// median of (first + 1, middle, last - 1) moved to the first position as pivot and an
// unguarded Hoare partition of [first + 1, last), recursing into the right part and looping
// on the left, down to ranges of 16 elements; heapsort for a range once the depth budget
// 2 floor(log2 n) is spent; then a straight insertion sort of the first 16 elements and an
// unguarded insertion of every later one.  Elements move as (key, value) pairs.
//
// Host and device code: kernels call it on LDS arrays, and a CPU test compares its
// permutations with std::sort's.
#ifndef NLI_INTROSORT_H_
#define NLI_INTROSORT_H_

#if defined(__HIPCC__)
#define NLI_HD __host__ __device__
#else
#define NLI_HD
#endif

namespace nli {

struct PairArrays {
    float *k;   // keys (sorted by)
    float *v;   // values (carried)
    NLI_HD bool less(int a, int b) const { return k[a] < k[b]; }
    NLI_HD void swap(int a, int b) const
    {
        const float tk = k[a], tv = v[a];
        k[a] = k[b];
        v[a] = v[b];
        k[b] = tk;
        v[b] = tv;
    }
    NLI_HD void copy(int dst, int src) const
    {
        k[dst] = k[src];
        v[dst] = v[src];
    }
};

// heap helpers over [first, first + len): __push_heap / __adjust_heap / __make_heap
NLI_HD inline void heap_push(const PairArrays &a, int first, int hole, int top, float vk, float vv)
{
    int parent = (hole - 1) / 2;
    while (hole > top && a.k[first + parent] < vk) {
        a.copy(first + hole, first + parent);
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a.k[first + hole] = vk;
    a.v[first + hole] = vv;
}

NLI_HD inline void heap_adjust(const PairArrays &a, int first, int hole, int len, float vk, float vv)
{
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (a.less(first + child, first + child - 1)) child--;
        a.copy(first + hole, first + child);
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        a.copy(first + hole, first + child - 1);
        hole = child - 1;
    }
    heap_push(a, first, hole, top, vk, vv);
}

// __partial_sort(first, last, last): make_heap, then sort_heap
NLI_HD inline void heap_sort(const PairArrays &a, int first, int last)
{
    const int len = last - first;
    if (len >= 2) {
        for (int parent = (len - 2) / 2;; parent--) {
            heap_adjust(a, first, parent, len, a.k[first + parent], a.v[first + parent]);
            if (parent == 0) break;
        }
    }
    for (int end = last; end - first > 1;) {
        end--;
        const float vk = a.k[end], vv = a.v[end];   // __pop_heap(first, end, end)
        a.copy(end, first);
        heap_adjust(a, first, 0, end - first, vk, vv);
    }
}

// __unguarded_linear_insert at position i; the `j > lo` guard never decides where the
// algorithm's invariant holds (a smaller-or-equal element lies to the left), it only
// bounds the loop
NLI_HD inline void linear_insert(const PairArrays &a, int lo, int i)
{
    const float vk = a.k[i], vv = a.v[i];
    int j = i;
    while (j > lo && vk < a.k[j - 1]) {
        a.copy(j, j - 1);
        j--;
    }
    a.k[j] = vk;
    a.v[j] = vv;
}

// __insertion_sort over [first, last)
NLI_HD inline void insertion_sort(const PairArrays &a, int first, int last)
{
    for (int i = first + 1; i < last; i++) {
        if (a.less(i, first)) {   // move_backward(first, i, i + 1), then *first = val
            const float vk = a.k[i], vv = a.v[i];
            for (int j = i; j > first; j--) a.copy(j, j - 1);
            a.k[first] = vk;
            a.v[first] = vv;
        } else {
            linear_insert(a, first, i);
        }
    }
}

// ranges pending on the explicit stack: at most the depth budget (16 for n <= 256) + 1
constexpr int kIntrosortStack = 20;

// std::sort(k, k + n) by key, carrying v; n <= 256; stk: 3 * kIntrosortStack ints of
// scratch (LDS in kernels)
NLI_HD inline void introsort_pairs(float *k, float *v, int n, int *stk)
{
    if (n <= 1) return;
    const PairArrays a{k, v};
    constexpr int kThreshold = 16;
    int lg = 0;
    while ((2 << lg) <= n) lg++;   // floor(log2 n)
    // __introsort_loop with an explicit stack of (first, last, depth budget); the ranges
    // on the stack are disjoint, so the order in which they are sorted does not matter
    int *st_first = stk, *st_last = stk + kIntrosortStack, *st_depth = stk + 2 * kIntrosortStack;
    int sp = 0;
    st_first[0] = 0;
    st_last[0] = n;
    st_depth[0] = 2 * lg;
    sp = 1;
    while (sp > 0) {
        sp--;
        int first = st_first[sp], last = st_last[sp], depth = st_depth[sp];
        while (last - first > kThreshold) {
            if (depth == 0) {
                heap_sort(a, first, last);
                break;
            }
            depth--;
            // __unguarded_partition_pivot: median of (first + 1, mid, last - 1) to first
            const int mid = first + (last - first) / 2;
            const int x = first + 1, y = mid, z = last - 1;
            if (a.less(x, y)) {
                if (a.less(y, z)) a.swap(first, y);
                else if (a.less(x, z)) a.swap(first, z);
                else a.swap(first, x);
            } else if (a.less(x, z)) {
                a.swap(first, x);
            } else if (a.less(y, z)) {
                a.swap(first, z);
            } else {
                a.swap(first, y);
            }
            // __unguarded_partition(first + 1, last, pivot = first)
            int lo = first + 1, hi = last;
            for (;;) {
                while (a.less(lo, first)) lo++;
                hi--;
                while (a.less(first, hi)) hi--;
                if (!(lo < hi)) break;
                a.swap(lo, hi);
                lo++;
            }
            // recurse into [cut, last), continue with [first, cut); each pending range
            // was pushed at a smaller depth budget, so sp <= 2 floor(log2 n) < kIntrosortStack
            st_first[sp] = lo;
            st_last[sp] = last;
            st_depth[sp] = depth;
            sp++;
            last = lo;
        }
    }
    // __final_insertion_sort
    if (n > kThreshold) {
        insertion_sort(a, 0, kThreshold);
        for (int i = kThreshold; i < n; i++) linear_insert(a, 0, i);
    } else {
        insertion_sort(a, 0, n);
    }
}

NLI_HD inline void introsort_pairs(float *k, float *v, int n)
{
    int stk[3 * kIntrosortStack];
    introsort_pairs(k, v, n, stk);
}

// The weighted median of patch_non_local_medians.cpp:113-125 walks the pairs sorted by
// colour and subtracts the weights from the float total tw until it drops below tw / 2;
// the order of equal colours changes only the rounding of that running total.  Given the
// pairs sorted by colour in any order of equal colours (k ascending), this tells whether
// every such order gives the same median: in exact arithmetic the total after each group
// of equal colours does not depend on the order, and a float total after j subtractions
// (each rounding to the nearest float, never above tw) lies within j * tw * 2^-24 of it.
// If the group c whose end brings the total below tw / 2 for every order is also the first
// group whose end can do so at all, every order crosses inside group c, whose colour is
// the median.  Sums are formed in double (exact up to j * tw * 2^-53).
NLI_HD inline bool median_order_free(const float *k, const float *w, int n, float tw)
{
    if (!(tw > 0.0f) || n <= 0) return false;
    const double half = (double)(tw / 2);
    const double unit = (double)tw * (0x1p-24 + 0x1p-50);
    double S = (double)tw;
    for (int e = 0; e < n;) {
        int j = e;
        while (j < n && k[j] == k[e]) S -= (double)w[j++];
        if (S - j * unit < half) return S + j * unit < half;   // first group that may cross
        e = j;
    }
    return false;   // no group crosses surely: the median depends on rounding
}

}  // namespace nli

#endif
