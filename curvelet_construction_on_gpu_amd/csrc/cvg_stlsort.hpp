// Tie-exact replay of libstdc++'s std::sort (GCC 11.4, bits/stl_algo.h and
// bits/stl_heap.h), usable on the host and on the device, for any element
// type and comparator: the neighbour lists' (distance, id) candidates and the
// spatial index's permutation of edge ids by cell key.
//
// cpu_curvelet orders each anchor's neighbours with an unstable std::sort by
// squared distance (preprocess.hpp:172-176).  When two candidates have exactly
// equal distances their final order is whatever libstdc++'s introsort makes of
// the scan order, so the GPU path replays the same published algorithm:
// introsort with median-of-three pivot moved to the front, an unguarded Hoare
// partition, depth limit 2*floor(log2 n), heap sort once the limit is spent,
// and a final insertion sort with threshold 16.  Distinct distances need none
// of this (any sort gives the unique order); the caller uses it only for lists
// that contain ties.  The spatial index (preprocess.hpp:90-92) sorts the edge
// ids by their packed cell key the same way, which fixes the order of the
// edges inside one pixel cell; the host replays it when that order matters.
#ifndef CVG_STLSORT_HPP
#define CVG_STLSORT_HPP

#include <hip/hip_runtime.h>

#define CVG_SORT_HD __host__ __device__ inline

namespace cvg {

struct Cand {
    double d;
    int id;
};

// The comparator of the neighbour sort, preprocess.hpp:172-176.
struct CandLess {
    CVG_SORT_HD bool operator()(const Cand &a, const Cand &b) const { return a.d < b.d; }
};

template <class T>
CVG_SORT_HD void elem_swap(T *a, T *b)
{
    const T t = *a;
    *a = *b;
    *b = t;
}

CVG_SORT_HD int floor_log2(int n)
{
    int l = 0;
    while (n > 1) {
        n >>= 1;
        ++l;
    }
    return l;
}

// __push_heap
template <class T, class Less>
CVG_SORT_HD void heap_push(T *first, int hole, int top, T value, const Less &less)
{
    int parent = (hole - 1) / 2;
    while (hole > top && less(first[parent], value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

// __adjust_heap
template <class T, class Less>
CVG_SORT_HD void heap_adjust(T *first, int hole, int len, T value, const Less &less)
{
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (less(first[child], first[child - 1])) child--;
        first[hole] = first[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        first[hole] = first[child - 1];
        hole = child - 1;
    }
    heap_push(first, hole, top, value, less);
}

// __make_heap followed by __sort_heap (what __partial_sort(first, last, last) does)
template <class T, class Less>
CVG_SORT_HD void heap_sort_range(T *first, int len, const Less &less)
{
    if (len >= 2) {
        int parent = (len - 2) / 2;
        while (true) {
            heap_adjust(first, parent, len, first[parent], less);
            if (parent == 0) break;
            parent--;
        }
    }
    for (int last = len; last > 1;) {
        --last;
        const T v = first[last];
        first[last] = first[0];
        heap_adjust(first, 0, last, v, less);
    }
}

// __move_median_to_first
template <class T, class Less>
CVG_SORT_HD void median_to_first(T *result, T *a, T *b, T *c, const Less &less)
{
    if (less(*a, *b)) {
        if (less(*b, *c))
            elem_swap(result, b);
        else if (less(*a, *c))
            elem_swap(result, c);
        else
            elem_swap(result, a);
    } else if (less(*a, *c))
        elem_swap(result, a);
    else if (less(*b, *c))
        elem_swap(result, c);
    else
        elem_swap(result, b);
}

// __unguarded_partition
template <class T, class Less>
CVG_SORT_HD T *unguarded_partition(T *first, T *last, const T *pivot, const Less &less)
{
    while (true) {
        while (less(*first, *pivot)) ++first;
        --last;
        while (less(*pivot, *last)) --last;
        if (!(first < last)) return first;
        elem_swap(first, last);
        ++first;
    }
}

// __unguarded_linear_insert
template <class T, class Less>
CVG_SORT_HD void linear_insert(T *last, const Less &less)
{
    const T val = *last;
    T *next = last - 1;
    while (less(val, *next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = val;
}

// __insertion_sort
template <class T, class Less>
CVG_SORT_HD void insertion_sort(T *first, T *last, const Less &less)
{
    if (first == last) return;
    for (T *i = first + 1; i != last; ++i) {
        if (less(*i, *first)) {
            const T val = *i;
            for (T *m = i; m != first; --m) *m = *(m - 1);
            *first = val;
        } else {
            linear_insert(i, less);
        }
    }
}

// std::sort(first, first + n, less).  The recursion of __introsort_loop is
// replaced by an explicit stack: segments are disjoint and each one's
// processing depends only on its own contents and depth budget, so the
// visiting order does not change the result.
template <class T, class Less>
CVG_SORT_HD void stl_sort(T *first, int n, const Less &less)
{
    if (n <= 0) return;
    const int threshold = 16;
    struct Seg {
        int lo, hi, depth;
    };
    Seg stack[64];  // at most depth limit 2*floor(log2 n) + 1 <= 61 entries
    int sp = 0;
    stack[sp++] = {0, n, 2 * floor_log2(n)};
    while (sp > 0) {
        Seg s = stack[--sp];
        while (s.hi - s.lo > threshold) {
            if (s.depth == 0) {
                heap_sort_range(first + s.lo, s.hi - s.lo, less);
                break;
            }
            --s.depth;
            T *f = first + s.lo;
            T *l = first + s.hi;
            T *mid = f + (l - f) / 2;
            median_to_first(f, f + 1, mid, l - 1, less);
            T *cut = unguarded_partition(f + 1, l, f, less);
            const int c = static_cast<int>(cut - first);
            stack[sp++] = {c, s.hi, s.depth};
            s.hi = c;
        }
    }
    // __final_insertion_sort
    if (n > threshold) {
        insertion_sort(first, first + threshold, less);
        for (T *i = first + threshold; i != first + n; ++i) linear_insert(i, less);
    } else {
        insertion_sort(first, first + n, less);
    }
}

// The neighbour sort by squared distance (preprocess.hpp:172-176).
CVG_SORT_HD void stl_sort(Cand *first, int n) { stl_sort(first, n, CandLess{}); }

}  // namespace cvg

#endif
