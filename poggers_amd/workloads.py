"""Synthetic workloads on the device (key streams are splitmix64 counters, so
every table's keys are distinct and reproducible from the seed)."""
from __future__ import annotations

import numpy as np
import torch

from . import gen_keys


def keys_at(seed: int, idx: np.ndarray) -> np.ndarray:
    """Host copy of the device key stream: splitmix64 output for counters
    `idx` of stream `seed` (the same function as pg_gen_keys)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    z[z == 0] = np.uint64(0x8000000000000000)
    z[z == np.uint64(0xFFFFFFFFFFFFFFFF)] = np.uint64(0x7FFFFFFFFFFFFFFF)
    return z


def log_uniform_sizes(ntables: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """Slot counts log-uniform in [2^lo, 2^hi]."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(2.0 ** lo), np.log(2.0 ** hi), ntables)).astype(np.uint64)


def batched_workload(sizes: np.ndarray, load: float, seed: int, device="cuda"):
    """Per table an insert batch filling it to `load` and a query batch of the
    same size, half present (the first half of its inserts) and half absent.
    Returns device tensors (ins_keys, ins_off, q_keys, q_off) and the host
    offsets; keys are int64 tensors holding uint64 bit patterns."""
    n = (np.floor(load * sizes.astype(np.float64))).astype(np.int64)
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(n)
    total = int(off[-1])
    ins = torch.empty(total, dtype=torch.int64, device=device)
    gen_keys(seed, 0, total, ins)
    npres = n // 2
    nabs = n - npres
    absent = torch.empty(int(nabs.sum()), dtype=torch.int64, device=device)
    gen_keys(seed, 1 << 40, absent.numel(), absent)
    q = torch.empty(total, dtype=torch.int64, device=device)
    a = 0
    for t in range(len(sizes)):
        s, p, m = int(off[t]), int(npres[t]), int(nabs[t])
        if p:
            q[s:s + p].copy_(ins[s:s + p])
        if m:
            q[s + p:s + p + m].copy_(absent[a:a + m])
        a += m
    d_off = torch.from_numpy(off.astype(np.int64)).to(device)
    return ins, d_off, q, d_off.clone(), off, npres


def absent_mask(off: np.ndarray, npres: np.ndarray, device="cuda"):
    """Bool tensor over a batched_workload query array: True for the absent
    keys (each table's queries after its first npres present ones)."""
    m = np.ones(int(off[-1]), dtype=bool)
    for t in range(len(npres)):
        m[int(off[t]):int(off[t]) + int(npres[t])] = False
    return torch.from_numpy(m).to(device)
