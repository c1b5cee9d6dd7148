// Tie-exact replay of libstdc++'s std::sort (GCC 11.4, bits/stl_algo.h and
// bits/stl_heap.h) for (distance, id) candidate lists, usable on the host and
// on the device.
//
// cpu_curvelet orders each anchor's neighbours with an unstable std::sort by
// squared distance (preprocess.hpp:172-176).  When two candidates have exactly
// equal distances their final order is whatever libstdc++'s introsort makes of
// the scan order, so the GPU path replays the same published algorithm:
// introsort with median-of-three pivot moved to the front, an unguarded Hoare
// partition, depth limit 2*floor(log2 n), heap sort once the limit is spent,
// and a final insertion sort with threshold 16.  Distinct distances need none
// of this (any sort gives the unique order); the caller uses it only for lists
// that contain ties.
#ifndef CVG_STLSORT_HPP
#define CVG_STLSORT_HPP

#include <hip/hip_runtime.h>

#define CVG_SORT_HD __host__ __device__ inline

namespace cvg {

struct Cand {
    double d;
    int id;
};

CVG_SORT_HD bool cand_less(const Cand &a, const Cand &b) { return a.d < b.d; }

CVG_SORT_HD void cand_swap(Cand *a, Cand *b)
{
    const Cand t = *a;
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
CVG_SORT_HD void heap_push(Cand *first, int hole, int top, Cand value)
{
    int parent = (hole - 1) / 2;
    while (hole > top && cand_less(first[parent], value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

// __adjust_heap
CVG_SORT_HD void heap_adjust(Cand *first, int hole, int len, Cand value)
{
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (cand_less(first[child], first[child - 1])) child--;
        first[hole] = first[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        first[hole] = first[child - 1];
        hole = child - 1;
    }
    heap_push(first, hole, top, value);
}

// __make_heap followed by __sort_heap (what __partial_sort(first, last, last) does)
CVG_SORT_HD void heap_sort_range(Cand *first, int len)
{
    if (len >= 2) {
        int parent = (len - 2) / 2;
        while (true) {
            heap_adjust(first, parent, len, first[parent]);
            if (parent == 0) break;
            parent--;
        }
    }
    for (int last = len; last > 1;) {
        --last;
        const Cand v = first[last];
        first[last] = first[0];
        heap_adjust(first, 0, last, v);
    }
}

// __move_median_to_first
CVG_SORT_HD void median_to_first(Cand *result, Cand *a, Cand *b, Cand *c)
{
    if (cand_less(*a, *b)) {
        if (cand_less(*b, *c))
            cand_swap(result, b);
        else if (cand_less(*a, *c))
            cand_swap(result, c);
        else
            cand_swap(result, a);
    } else if (cand_less(*a, *c))
        cand_swap(result, a);
    else if (cand_less(*b, *c))
        cand_swap(result, c);
    else
        cand_swap(result, b);
}

// __unguarded_partition
CVG_SORT_HD Cand *unguarded_partition(Cand *first, Cand *last, const Cand *pivot)
{
    while (true) {
        while (cand_less(*first, *pivot)) ++first;
        --last;
        while (cand_less(*pivot, *last)) --last;
        if (!(first < last)) return first;
        cand_swap(first, last);
        ++first;
    }
}

// __unguarded_linear_insert
CVG_SORT_HD void linear_insert(Cand *last)
{
    const Cand val = *last;
    Cand *next = last - 1;
    while (cand_less(val, *next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = val;
}

// __insertion_sort
CVG_SORT_HD void insertion_sort(Cand *first, Cand *last)
{
    if (first == last) return;
    for (Cand *i = first + 1; i != last; ++i) {
        if (cand_less(*i, *first)) {
            const Cand val = *i;
            for (Cand *m = i; m != first; --m) *m = *(m - 1);
            *first = val;
        } else {
            linear_insert(i);
        }
    }
}

// std::sort(first, first + n, [](a, b) { return a.d < b.d; }).  The recursion
// of __introsort_loop is replaced by an explicit stack: segments are disjoint
// and each one's processing depends only on its own contents and depth budget,
// so the visiting order does not change the result.
CVG_SORT_HD void stl_sort(Cand *first, int n)
{
    if (n <= 0) return;
    const int threshold = 16;
    struct Seg {
        int lo, hi, depth;
    };
    Seg stack[48];  // depth limit 2*floor(log2 n) + 1 entries; n < 2^23
    int sp = 0;
    stack[sp++] = {0, n, 2 * floor_log2(n)};
    while (sp > 0) {
        Seg s = stack[--sp];
        while (s.hi - s.lo > threshold) {
            if (s.depth == 0) {
                heap_sort_range(first + s.lo, s.hi - s.lo);
                break;
            }
            --s.depth;
            Cand *f = first + s.lo;
            Cand *l = first + s.hi;
            Cand *mid = f + (l - f) / 2;
            median_to_first(f, f + 1, mid, l - 1);
            Cand *cut = unguarded_partition(f + 1, l, f);
            const int c = static_cast<int>(cut - first);
            stack[sp++] = {c, s.hi, s.depth};
            s.hi = c;
        }
    }
    // __final_insertion_sort
    if (n > threshold) {
        insertion_sort(first, first + threshold);
        for (Cand *i = first + threshold; i != first + n; ++i) linear_insert(i);
    } else {
        insertion_sort(first, first + n);
    }
}

}  // namespace cvg

#endif
