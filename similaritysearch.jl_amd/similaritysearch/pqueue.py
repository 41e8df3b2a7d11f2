"""k-NN result queues of the reference (src/pqueue/), host side.

The queue-vector form of searchbatch! (src/SimilaritySearch.jl:384-398) hands one
queue per query to the index; the GPU computes each query's first k results in
(distance, id) order and this module fills the caller's queues with the reference's own
push semantics:

- KnnSorted (src/pqueue/knnsorted.jl:31-233): ids/dists kept sorted ascending; a push
  into a full queue is rejected when d >= maximum, else replaces the last entry, and the
  new entry is placed after every entry of equal distance (stable).
- KnnHeap (src/pqueue/knnheap.jl:29-187): a binary max-heap on distance (root = farthest)
  with the nearest tracked apart; a push into a full heap is rejected when d >= maximum,
  else replaces the root and sifts down.  sortitems heapsorts the active items.

Both keep the same set of items for the same pushes unless distances tie at the k-th
position, where the heap may keep another of the tied items.
"""
import numpy as np


class _Queue:
    def __init__(self, ids, dists):
        ids = np.asarray(ids)
        dists = np.asarray(dists)
        if ids.dtype != np.uint32 or dists.dtype != np.float32 or ids.shape != dists.shape:
            raise TypeError("a queue takes UInt32 ids and Float32 dists of equal length")
        self.ids, self.dists = ids, dists
        self.maxlen = len(ids)
        self.len = 0

    def __len__(self):
        return self.len

    def maxlength(self):
        return self.maxlen


class KnnSorted(_Queue):
    """KnnSorted: active items ids[sp:ep+1], sorted ascending by distance."""

    def __init__(self, ids, dists):
        super().__init__(ids, dists)
        self.sp = 0

    def reuse(self, maxlen=None):
        """reuse! (knnsorted.jl:206-216)."""
        self.maxlen = len(self.ids) if maxlen is None else int(maxlen)
        self.sp, self.len = 0, 0
        return self

    def push_item(self, i, d):
        """push_item! (knnsorted.jl:140-168); returns whether the item was kept."""
        d = np.float32(d)
        ids, dists = self.ids, self.dists
        if self.len < self.maxlen:
            if self.sp + self.len == len(ids):   # compaction (:145-153)
                ids[:self.len] = ids[self.sp:self.sp + self.len].copy()
                dists[:self.len] = dists[self.sp:self.sp + self.len].copy()
                self.sp = 0
            ep = self.sp + self.len
            self.len += 1
        else:
            ep = self.sp + self.len - 1
            if d >= dists[ep]:
                return False
        # sort_last_item! (:68-101): first slot whose distance is > d
        lo = self.sp + int(np.searchsorted(dists[self.sp:ep], d, side="right"))
        ids[lo + 1:ep + 1] = ids[lo:ep].copy()
        dists[lo + 1:ep + 1] = dists[lo:ep].copy()
        ids[lo], dists[lo] = np.uint32(i), d
        return True

    def pop_min(self):
        """pop_min! (knnsorted.jl:191-196)."""
        i, d = self.ids[self.sp], self.dists[self.sp]
        self.sp += 1
        self.len -= 1
        return int(i), float(d)

    def nearest(self):
        return int(self.ids[self.sp]), float(self.dists[self.sp])

    def maximum(self):
        return float(self.dists[self.sp + self.len - 1])

    def id_view(self):
        return self.ids[self.sp:self.sp + self.len]

    def dist_view(self):
        return self.dists[self.sp:self.sp + self.len]

    def sortitems(self):
        return self.id_view(), self.dist_view()


class KnnHeap(_Queue):
    """KnnHeap: a max-heap in ids[:len] / dists[:len] (1-based heap positions)."""

    def __init__(self, ids, dists):
        super().__init__(ids, dists)
        self.min_id, self.min_dist = 0, np.float32(np.inf)

    def reuse(self, maxlen=None):
        """reuse! (knnheap.jl:168-175)."""
        self.maxlen = len(self.ids) if maxlen is None else int(maxlen)
        self.len = 0
        self.min_id, self.min_dist = 0, np.float32(np.inf)
        return self

    def _swap(self, a, b):
        I, D = self.ids, self.dists
        I[a - 1], I[b - 1] = I[b - 1], I[a - 1]
        D[a - 1], D[b - 1] = D[b - 1], D[a - 1]

    def _lt(self, a, b):
        return self.dists[a - 1] < self.dists[b - 1]

    def _fix_up(self, i):
        """heapfix_up! (heap.jl:17-27)."""
        while (p := i >> 1) > 0 and self._lt(p, i):
            self._swap(i, p)
            i = p

    def _fix_down(self, n):
        """heapfix_down! (heap.jl:36-50)."""
        i = 1
        while (l := 2 * i) <= n:
            r = l + 1
            if r > n or self._lt(r, l):
                if self._lt(l, i):
                    break
                self._swap(i, l)
                i = l
            else:
                if self._lt(r, i):
                    break
                self._swap(i, r)
                i = r

    def push_item(self, i, d):
        """push_item! (knnheap.jl:102-127); returns whether the item was kept."""
        d = np.float32(d)
        if self.len < self.maxlen:
            self.len += 1
            self.ids[self.len - 1], self.dists[self.len - 1] = np.uint32(i), d
            self._fix_up(self.len)
            if self.len == 1 or d < self.min_dist:
                self.min_id, self.min_dist = int(i), d
            return True
        if d >= self.dists[0]:
            return False
        self.ids[0], self.dists[0] = np.uint32(i), d
        self._fix_down(self.len)
        if d < self.min_dist:
            self.min_id, self.min_dist = int(i), d
        return True

    def nearest(self):
        return self.min_id, float(self.min_dist)

    def maximum(self):
        return float(self.dists[0])

    def id_view(self):
        return self.ids[:self.len]

    def dist_view(self):
        return self.dists[:self.len]

    def sortitems(self):
        """sortitems! (knnheap.jl:94-97): heapsort of the active items, ascending; the heap
        structure is gone afterwards."""
        for i in range(self.len, 1, -1):
            self._swap(1, i)
            self._fix_down(i - 1)
        return self.id_view(), self.dist_view()


def knnqueue(T, k):
    """knnqueue(T, k) (src/pqueue/pqueue.jl:229-247): a queue of capacity k."""
    if T not in (KnnSorted, KnnHeap):
        raise TypeError("knnqueue takes KnnSorted or KnnHeap")
    return T(np.zeros(int(k), np.uint32), np.zeros(int(k), np.float32))


def fill(res, ids, dists):
    """Pushes a result run (ascending, 0 ends it) into res as it stands: the reference's
    queue-vector searchbatch! pushes into the given queues without emptying them
    (src/SimilaritySearch.jl:384-398).  The GPU search itself starts empty, so a queue
    that already holds items gets this search's results merged in."""
    for i, d in zip(ids, dists):
        if i == 0:
            break
        res.push_item(int(i), d)
    return res
