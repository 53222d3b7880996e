"""Sort-based numbering on the device for the cold path (``nfx_first_appearance``,
``nfx_stable_sort64``).

The reference numbers mesh entities as its builder meets them (``src/mesh/mesh.py:239-339``),
orders tag-filtered entities by first appearance (``src/mesh/mesh_topology.py:28-35``) and
forms the matrix pattern as the union of element entries (``setValue(addv=True)``,
``src/assembly/ScatterFormData.py:5-11``). Each of these reduces to a stable sort of int64
keys; with ``device="cuda"`` the mesh builder, the topology and the pattern build run that
sort on the GPU (radix sort + scans) and keep the O(K) glue in numpy. The results are the
same integers as the host path, which the GPU tests check bit for bit.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from .device import torch


def _key_bits(keys):
    m = int(keys.max()) if keys.size else 0
    if m < 0 or (keys.size and int(keys.min()) < 0):
        raise ValueError("keys must be non-negative")
    return max(1, m.bit_length())


def _workspace(K, device):
    t = torch()
    nb = ctypes.c_size_t()
    nat.call("nfx_sort_workspace", int(K), ctypes.byref(nb))
    return t.empty(max(int(nb.value), 1), dtype=t.uint8, device=device), int(nb.value)


def first_appearance_ids(keys, device="cuda"):
    """``(ids, first)`` of the host helper: ids (K,) int64 ranks by first appearance,
    first (n,) int64 input position of each distinct key's first occurrence."""
    t = torch()
    keys = np.ascontiguousarray(np.asarray(keys, dtype=np.int64).reshape(-1))
    K = len(keys)
    if K == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    d_keys = t.from_numpy(keys).to(device)
    ids = t.empty(K, dtype=t.int32, device=device)
    first = t.empty(K, dtype=t.int32, device=device)
    count = t.empty(1, dtype=t.int32, device=device)
    work, nbytes = _workspace(K, device)
    s = nat.stream_handle()
    nat.call("nfx_first_appearance", K, _key_bits(keys), nat.ptr(d_keys), nat.ptr(ids), nat.ptr(first),
             nat.ptr(count), nat.ptr(work), nbytes, s)
    n = int(count.item())
    return ids.cpu().numpy().astype(np.int64), first[:n].cpu().numpy().astype(np.int64)


def stable_argsort(keys, device="cuda"):
    """``np.argsort(keys, kind="stable")`` and the sorted keys, sorted on the device."""
    t = torch()
    keys = np.ascontiguousarray(np.asarray(keys, dtype=np.int64).reshape(-1))
    K = len(keys)
    if K == 0:
        return np.zeros(0, np.int64), keys.copy()
    d_keys = t.from_numpy(keys).to(device)
    sk = t.empty(K, dtype=t.int64, device=device)
    perm = t.empty(K, dtype=t.int32, device=device)
    work, nbytes = _workspace(K, device)
    nat.call("nfx_stable_sort64", K, _key_bits(keys), nat.ptr(d_keys), nat.ptr(sk), nat.ptr(perm), nat.ptr(work),
             nbytes, nat.stream_handle())
    return perm.cpu().numpy().astype(np.int64), sk.cpu().numpy()
