"""Byte-bounded LRU of device-resident matrices and factors.

The solve server keeps the assembled FP64 matrix and the factors of each
rigid-body geometry here, and the free swimmer keeps the FP32 factors of its
body-frame matrix for every body shape it has met (a periodic beat revisits
the same shapes), so repeated right-hand sides and repeated shapes cost only
triangular solves and refinement.
"""
from __future__ import annotations

import threading
from collections import OrderedDict


class FactorCache:
    """LRU of resident items per key, bounded in bytes (thread-safe)."""

    def __init__(self, max_bytes: int):
        self.max_bytes = int(max_bytes)
        self.bytes = 0
        self.items: OrderedDict[str, tuple] = OrderedDict()
        self.lock = threading.Lock()
        self.hits = 0
        self.misses = 0

    def get(self, key):
        with self.lock:
            item = self.items.get(key)
            if item is not None:
                self.items.move_to_end(key)
                self.hits += 1
            else:
                self.misses += 1
            return item

    def put(self, key, item, nbytes):
        """Store `item` (a tuple) with its size; the stored tuple gains the size
        as its last element.  Items larger than the whole budget are not kept."""
        with self.lock:
            if nbytes > self.max_bytes:
                return
            old = self.items.pop(key, None)
            if old is not None:
                self.bytes -= old[-1]
            while self.bytes + nbytes > self.max_bytes and self.items:
                _, ev = self.items.popitem(last=False)
                self.bytes -= ev[-1]
            self.items[key] = (*item, nbytes)
            self.bytes += nbytes
