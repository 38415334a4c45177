"""Byte-bounded LRU of device-resident matrices and factors.

The solve server keeps the assembled FP64 matrix and the factors of each
rigid-body geometry here, and the free swimmer keeps the FP32 inverse of its
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
        self.inflight: dict = {}   # key (or (key, "upgrade")) -> Event of the thread making it
        self.key_uses: dict = {}   # counted lookups (hits and makes) per stored key

    def get_or_make(self, key, make, made: list | None = None, count: bool = True):
        """The item for `key`, made once: on a miss ``make()`` returns
        ``(item, nbytes)`` and is stored; a caller that finds another thread
        making the same key waits for it instead of making it again (and
        makes it itself if that attempt failed or the item was not kept).
        ``make`` must leave the item complete (its device work finished)
        before returning, since callers use it on their own streams.  With
        ``made`` (a list), True is appended when this call made the item.
        count=False leaves the hit statistics alone (a prefetching caller);
        an item it makes still counts as a miss."""
        while True:
            with self.lock:
                item = self.items.get(key)
                if item is not None:
                    self.items.move_to_end(key)
                    if count:
                        self.hits += 1
                        self.key_uses[key] = self.key_uses.get(key, 0) + 1
                    if made is not None:
                        made.append(False)
                    return item
                ev = self.inflight.get(key)
                if ev is None:
                    self.misses += 1
                    ev = threading.Event()
                    self.inflight[key] = ev
                    break
            ev.wait()   # then look again: stored (a hit), or the maker failed (make it here)
        try:
            item, nbytes = make()
            self.put(key, item, nbytes)
            if made is not None:
                made.append(True)
            with self.lock:
                if count and key in self.items:
                    self.key_uses[key] = self.key_uses.get(key, 0) + 1
                stored = self.items.get(key)   # the stored tuple itself, for upgrade()'s identity check
            same = stored is not None and len(stored) == len(item) + 1 and all(a is b for a, b in zip(stored, item))
            return stored if same else (*item, nbytes)
        finally:
            with self.lock:
                self.inflight.pop(key, None)
            ev.set()

    def upgrade(self, key, old, make):
        """Replace the stored item `old` (as returned for `key`) by
        ``make()`` -> ``(item, nbytes)``, once: if another thread is already
        upgrading this key, or the stored item is no longer `old`, return the
        current item (or `old` if the key is gone) without waiting.  ``make``
        must leave its item complete before returning."""
        tag = (key, "upgrade")
        with self.lock:
            cur = self.items.get(key)
            if cur is not old or tag in self.inflight:
                return cur if cur is not None else old
            ev = threading.Event()
            self.inflight[tag] = ev
        try:
            item, nbytes = make()
            with self.lock:
                keep = self.items.get(key) is old
            if keep:
                self.put(key, item, nbytes)
            return (*item, nbytes)
        finally:
            with self.lock:
                self.inflight.pop(tag, None)
            ev.set()

    def clear(self) -> None:
        """Drop every item (their device memory goes back to the allocator);
        the hit and miss counts stay."""
        with self.lock:
            self.items.clear()
            self.key_uses.clear()
            self.bytes = 0

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
                gone, ev = self.items.popitem(last=False)
                self.bytes -= ev[-1]
                self.key_uses.pop(gone, None)
            self.items[key] = (*item, nbytes)
            self.bytes += nbytes
