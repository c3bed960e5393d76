"""std::sort and std::stable_sort of GNU libstdc++ 11 (GCC 11.4), restated in
Python (test infrastructure).

The reference orders the references it prints with these two library calls
(MirAge/Analyser.cpp:404 and :266-293), and the printed order is whatever
permutation the library's algorithm produces: std::sort is not stable, and the
taxonomy comparator of :266-293 is not a strict weak ordering.  The published
algorithm of libstdc++ 11 (bits/stl_algo.h, bits/stl_heap.h), as this image's
g++ 11.4 and hipcc build it:

* std::sort: introsort with depth limit 2 * floor(log2(n)); partitions with the
  median of (first + 1, middle, last - 1) moved to the front, unguarded
  partition, and stops at 16 elements; heap sort when the depth limit is hit;
  then insertion sort (the first 16 guarded, the rest unguarded).
* std::stable_sort with its temporary buffer of ceil(n / 2) elements: each half
  is sorted by insertion sort on chunks of 7 and merge passes through the
  buffer, then the halves are merged through the buffer (forward when the
  first half is not longer, backward otherwise).

comp(a, b) is the library's "a goes before b" predicate.  Pinned against the
library itself on random inputs (tests/test_results.py) and, end to end, by the
result logs the reference wrote.
"""


def _move_median_to_first(a, result, i, j, k, comp):
    if comp(a[i], a[j]):
        if comp(a[j], a[k]):
            a[result], a[j] = a[j], a[result]
        elif comp(a[i], a[k]):
            a[result], a[k] = a[k], a[result]
        else:
            a[result], a[i] = a[i], a[result]
    elif comp(a[i], a[k]):
        a[result], a[i] = a[i], a[result]
    elif comp(a[j], a[k]):
        a[result], a[k] = a[k], a[result]
    else:
        a[result], a[j] = a[j], a[result]


def _unguarded_partition(a, first, last, pivot, comp):
    while True:
        while comp(a[first], a[pivot]):
            first += 1
        last -= 1
        while comp(a[pivot], a[last]):
            last -= 1
        if not first < last:
            return first
        a[first], a[last] = a[last], a[first]
        first += 1


def _push_heap(a, base, hole, top, value, comp):
    parent = (hole - 1) // 2
    while hole > top and comp(a[base + parent], value):
        a[base + hole] = a[base + parent]
        hole = parent
        parent = (hole - 1) // 2
    a[base + hole] = value


def _adjust_heap(a, base, hole, n, value, comp):
    top = hole
    child = hole
    while child < (n - 1) // 2:
        child = 2 * (child + 1)
        if comp(a[base + child], a[base + child - 1]):
            child -= 1
        a[base + hole] = a[base + child]
        hole = child
    if (n & 1) == 0 and child == (n - 2) // 2:
        child = 2 * (child + 1)
        a[base + hole] = a[base + child - 1]
        hole = child - 1
    _push_heap(a, base, hole, top, value, comp)


def _heap_sort(a, first, last, comp):
    n = last - first
    if n >= 2:  # make_heap
        parent = (n - 2) // 2
        while True:
            _adjust_heap(a, first, parent, n, a[first + parent], comp)
            if parent == 0:
                break
            parent -= 1
    while last - first > 1:  # sort_heap: pop_heap(first, last - 1, last - 1)
        last -= 1
        value = a[last]
        a[last] = a[first]
        _adjust_heap(a, first, 0, last - first, value, comp)


def _introsort_loop(a, first, last, depth, comp):
    while last - first > 16:
        if depth == 0:
            _heap_sort(a, first, last, comp)
            return
        depth -= 1
        mid = first + (last - first) // 2
        _move_median_to_first(a, first, first + 1, mid, last - 1, comp)
        cut = _unguarded_partition(a, first + 1, last, first, comp)
        _introsort_loop(a, cut, last, depth, comp)
        last = cut


def _unguarded_linear_insert(a, i, comp):
    val = a[i]
    j = i - 1
    while comp(val, a[j]):
        a[i] = a[j]
        i = j
        j -= 1
    a[i] = val


def _insertion_sort(a, first, last, comp):
    if first == last:
        return
    for i in range(first + 1, last):
        if comp(a[i], a[first]):
            val = a[i]
            a[first + 1:i + 1] = a[first:i]
            a[first] = val
        else:
            _unguarded_linear_insert(a, i, comp)


def sort(a, comp):
    """std::sort(a.begin(), a.end(), comp), in place."""
    n = len(a)
    if n == 0:
        return
    _introsort_loop(a, 0, n, 2 * (n.bit_length() - 1), comp)
    if n > 16:
        _insertion_sort(a, 0, 16, comp)
        for i in range(16, n):
            _unguarded_linear_insert(a, i, comp)
    else:
        _insertion_sort(a, 0, n, comp)


def _move_merge(src, f1, l1, f2, l2, dst, out, comp):
    while f1 != l1 and f2 != l2:
        if comp(src[f2], src[f1]):
            dst[out] = src[f2]
            f2 += 1
        else:
            dst[out] = src[f1]
            f1 += 1
        out += 1
    for x in range(f1, l1):
        dst[out] = src[x]
        out += 1
    for x in range(f2, l2):
        dst[out] = src[x]
        out += 1
    return out


def _merge_sort_loop(src, first, last, dst, out, step, comp):
    two = 2 * step
    while last - first >= two:
        out = _move_merge(src, first, first + step, first + step, first + two, dst, out, comp)
        first += two
    step = min(last - first, step)
    _move_merge(src, first, first + step, first + step, last, dst, out, comp)


def _merge_sort_with_buffer(a, first, last, buf, comp):
    n = last - first
    step = 7
    f = first
    while last - f >= step:  # chunk insertion sort
        _insertion_sort(a, f, f + step, comp)
        f += step
    _insertion_sort(a, f, last, comp)
    while step < n:
        _merge_sort_loop(a, first, last, buf, 0, step, comp)
        step *= 2
        _merge_sort_loop(buf, 0, n, a, first, step, comp)
        step *= 2


def stable_sort(a, comp):
    """std::stable_sort(a.begin(), a.end(), comp), in place, with the full
    temporary buffer the library obtains for small ranges."""
    n = len(a)
    if n == 0:
        return
    half = (n + 1) // 2
    buf = [None] * half
    middle = half
    _merge_sort_with_buffer(a, 0, middle, buf, comp)
    _merge_sort_with_buffer(a, middle, n, buf, comp)
    len1, len2 = middle, n - middle
    if len1 <= len2:  # forward merge from the buffer (__move_merge_adaptive)
        buf[:len1] = a[0:middle]
        f1, l1, f2, out = 0, len1, middle, 0
        while f1 != l1 and f2 != n:
            if comp(a[f2], buf[f1]):
                a[out] = a[f2]
                f2 += 1
            else:
                a[out] = buf[f1]
                f1 += 1
            out += 1
        while f1 != l1:
            a[out] = buf[f1]
            f1 += 1
            out += 1
    else:  # backward merge (__move_merge_adaptive_backward)
        buf[:len2] = a[middle:n]
        if len2 == 0:
            return
        l1 = middle - 1          # last of the first half, in a
        l2 = len2 - 1            # last of the second half, in buf
        res = n
        while True:
            if comp(buf[l2], a[l1]):
                res -= 1
                a[res] = a[l1]
                if l1 == 0:
                    for x in range(l2, -1, -1):
                        res -= 1
                        a[res] = buf[x]
                    return
                l1 -= 1
            else:
                res -= 1
                a[res] = buf[l2]
                if l2 == 0:
                    return
                l2 -= 1
