"""Canonical form of an allocation (cell -> cache position), independent of the
nondeterministic append order of the reference's atomicAdd."""
from __future__ import annotations

import numpy as np


def canonical(count: int, cav: np.ndarray, entries: np.ndarray, capacity: int):
    """Returns (cells, positions, phantoms, addr_ok).

    cells      sorted flat indices of claimed CAV cells;
    positions  uint32 bit patterns [n, 3] of Entry[CAV - 1].Position per cell
               (rows for addresses >= capacity are all 0xFFFFFFFF);
    phantoms   sorted list of position bit patterns of stored caches that no
               cell points to (corners outside the address volume);
    addr_ok    True if the CAV values minus one are distinct and < count.
    """
    flat = cav.reshape(-1)
    cells = np.flatnonzero(flat)
    addr = flat[cells].astype(np.int64) - 1
    live = min(count, capacity)
    addr_ok = bool(len(np.unique(addr)) == len(addr) and (len(addr) == 0 or addr.max() < count))
    pos = np.full((len(cells), 3), 0xFFFFFFFF, np.uint32)
    inb = addr < live
    pos[inb] = entries[addr[inb], :3].view(np.uint32)
    referenced = np.zeros(live, bool)
    referenced[addr[inb]] = True
    ph = entries[:live][~referenced, :3].view(np.uint32)
    phantoms = sorted(map(tuple, ph.tolist()))
    return cells, pos, phantoms, addr_ok


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
