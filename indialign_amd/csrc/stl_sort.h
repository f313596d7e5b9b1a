// Deterministic restatement of the GCC libstdc++ std::sort algorithm
// (introsort: median-of-three quicksort with an unguarded Hoare partition,
// heapsort once the depth budget 2*floor(log2 n) is spent, and a final
// insertion sort with threshold 16).
//
// Why it exists: the anchor-seed builder of the reference sorts
// (count, index) pairs by count only with std::sort (seeds.h:237-238), which
// is not stable; which of several equally-ranked anchors survive depends on
// the exact permutation that libstdc++ produces.  The device code needs that
// same permutation, so the algorithm's steps are kept alike here, in code of
// our own, and tests/test_stl_sort.py checks it against the host std::sort on
// random and tie-heavy inputs.
//
// Usable on host and device; operates in place on T* with a comparator
// object `less(a, b)`.
#pragma once

#ifndef IA_SORT_HD
#if defined(__HIPCC__)
#define IA_SORT_HD __host__ __device__ inline
#else
#define IA_SORT_HD inline
#endif
#endif

namespace ia_sort {

template <class T>
IA_SORT_HD void swap_(T &a, T &b) {
    T t = a;
    a = b;
    b = t;
}

template <class T, class L>
IA_SORT_HD void push_heap_(T *f, long hole, long top, T val, L less) {
    long parent = (hole - 1) / 2;
    while (hole > top && less(f[parent], val)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = val;
}

template <class T, class L>
IA_SORT_HD void adjust_heap_(T *f, long hole, long len, T val, L less) {
    const long top = hole;
    long child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (less(f[child], f[child - 1])) child--;
        f[hole] = f[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        f[hole] = f[child - 1];
        hole = child - 1;
    }
    push_heap_(f, hole, top, val, less);
}

template <class T, class L>
IA_SORT_HD void make_heap_(T *f, long len, L less) {
    if (len < 2) return;
    long parent = (len - 2) / 2;
    while (true) {
        T v = f[parent];
        adjust_heap_(f, parent, len, v, less);
        if (parent == 0) return;
        parent--;
    }
}

// partial_sort(first, last, last): heap select (a no-op scan when middle ==
// last) then sort_heap.
template <class T, class L>
IA_SORT_HD void heap_sort_(T *f, long len, L less) {
    make_heap_(f, len, less);
    while (len > 1) {
        len--;
        T v = f[len];
        f[len] = f[0];
        adjust_heap_(f, 0L, len, v, less);
    }
}

template <class T, class L>
IA_SORT_HD void median_to_first_(T *res, T *a, T *b, T *c, L less) {
    if (less(*a, *b)) {
        if (less(*b, *c)) swap_(*res, *b);
        else if (less(*a, *c)) swap_(*res, *c);
        else swap_(*res, *a);
    } else if (less(*a, *c)) {
        swap_(*res, *a);
    } else if (less(*b, *c)) {
        swap_(*res, *c);
    } else {
        swap_(*res, *b);
    }
}

template <class T, class L>
IA_SORT_HD T *unguarded_partition_(T *first, T *last, T *pivot, L less) {
    while (true) {
        while (less(*first, *pivot)) ++first;
        --last;
        while (less(*pivot, *last)) --last;
        if (!(first < last)) return first;
        swap_(*first, *last);
        ++first;
    }
}

template <class T, class L>
IA_SORT_HD void introsort_loop_(T *first, T *last, int depth, L less) {
    // Recursion on the right part is turned into an explicit stack; the
    // visiting order (right part fully, then continue on the left) is the
    // same as the recursive original, so the permutation is unchanged.
    T *stk_f[64], *stk_l[64];
    int stk_d[64];
    int sp = 0;
    stk_f[sp] = first; stk_l[sp] = last; stk_d[sp] = depth; sp++;
    while (sp > 0) {
        sp--;
        T *f = stk_f[sp], *l = stk_l[sp];
        int d = stk_d[sp];
        // body of the libstdc++ loop for one call frame; the frame's
        // remaining left part is pushed *below* the right-part call so it
        // runs after the right part completes.
        while (l - f > 16) {
            if (d == 0) {
                heap_sort_(f, (long)(l - f), less);
                break;
            }
            --d;
            T *mid = f + (l - f) / 2;
            median_to_first_(f, f + 1, mid, l - 1, less);
            T *cut = unguarded_partition_(f + 1, l, f, less);
            // emulate: introsort_loop(cut, l, d); l = cut; (continue loop)
            stk_f[sp] = f; stk_l[sp] = cut; stk_d[sp] = d; sp++;   // left continues later
            f = cut;                                                // right now
        }
    }
}

template <class T, class L>
IA_SORT_HD void unguarded_linear_insert_(T *last, L less) {
    T val = *last;
    T *next = last - 1;
    while (less(val, *next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = val;
}

template <class T, class L>
IA_SORT_HD void insertion_sort_(T *first, T *last, L less) {
    if (first == last) return;
    for (T *i = first + 1; i != last; ++i) {
        if (less(*i, *first)) {
            T val = *i;
            for (T *p = i; p != first; --p) *p = *(p - 1);
            *first = val;
        } else {
            unguarded_linear_insert_(i, less);
        }
    }
}

template <class T, class L>
IA_SORT_HD void sort(T *first, T *last, L less) {
    if (first == last) return;
    long n = (long)(last - first);
    int lg = 0;
    while ((n >> (lg + 1)) != 0) lg++;
    introsort_loop_(first, last, 2 * lg, less);
    if (last - first > 16) {
        insertion_sort_(first, first + 16, less);
        for (T *i = first + 16; i != last; ++i) unguarded_linear_insert_(i, less);
    } else {
        insertion_sort_(first, last, less);
    }
}

}  // namespace ia_sort
