"""Helpers shared by the GPU parity tests: build matching oracle / HIP tables
from one descriptor and move uint64 arrays between numpy and torch."""
import numpy as np
import torch

import oracle
import poggers_amd as pg


def dev_u64(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).cuda()


def dev_u16(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint16).view(np.int16)).cuda()


def dev_u8(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint8)).cuda()


def host(t: torch.Tensor, dtype) -> np.ndarray:
    return t.cpu().numpy().view(dtype)


def pair(desc: pg.TableDesc, seed: int = 42):
    """(oracle table, GPU table) built from the same descriptor."""
    o = oracle.Table(desc.rep, [lv.as_dict() for lv in desc.levels], seed=seed)
    g = pg.Table(desc, seed=seed)
    return o, g


class Ops:
    """Bulk ops on a GPU table with host arrays in and out."""

    def __init__(self, table: pg.Table):
        self.t = table
        self.kv = table.is_kv
        # value buffers hold the format's Val type: uint64 / uint16 / uint8
        self.narrow = table.desc.rep in (pg.REP_FP8_V8, pg.REP_FP10_V6)

    def _dev_vals(self, vals):
        if self.kv:
            return dev_u64(vals)
        return dev_u8(vals) if self.narrow else dev_u16(vals)

    def _out_vals(self, n):
        dt = torch.int64 if self.kv else (torch.uint8 if self.narrow else torch.int16)
        return torch.zeros(n, dtype=dt, device="cuda")

    def _host_vals(self, t):
        return host(t, np.uint64 if self.kv else (np.uint8 if self.narrow else np.uint16))

    def insert(self, keys, vals=None, with_delete=False):
        n = len(keys)
        k = dev_u64(keys)
        v = None
        if vals is not None:
            v = self._dev_vals(vals)
        st = torch.zeros(n, dtype=torch.uint8, device="cuda")
        fails = torch.zeros(1, dtype=torch.int64, device="cuda")
        self.t.bulk_insert(k, v, st, fails, with_delete=with_delete)
        torch.cuda.synchronize()
        return st.cpu().numpy(), int(fails.item())

    def query(self, keys):
        n = len(keys)
        k = dev_u64(keys)
        found = torch.zeros(n, dtype=torch.uint8, device="cuda")
        vals = self._out_vals(n)
        self.t.bulk_query(k, vals, found)
        torch.cuda.synchronize()
        return found.cpu().numpy(), self._host_vals(vals)

    def upsert(self, keys, vals=None, delete=False):
        n = len(keys)
        k = dev_u64(keys)
        v = None if vals is None else self._dev_vals(vals)
        st = torch.zeros(n, dtype=torch.uint8, device="cuda")
        found = torch.zeros(n, dtype=torch.uint8, device="cuda")
        out = self._out_vals(n)
        self.t.bulk_insert_if_not_exists(k, v, out, found, st, delete=delete)
        torch.cuda.synchronize()
        return st.cpu().numpy(), found.cpu().numpy(), self._host_vals(out)

    def remove(self, keys):
        n = len(keys)
        k = dev_u64(keys)
        found = torch.zeros(n, dtype=torch.uint8, device="cuda")
        self.t.bulk_remove(k, found)
        torch.cuda.synchronize()
        return found.cpu().numpy()

    def level(self, lvl):
        dt = {pg.REP_KV64: np.uint64, pg.REP_FP16_V16: np.uint32, pg.REP_FP8_V8: np.uint16,
              pg.REP_FP10_V6: np.uint16, pg.REP_FP12: np.uint16, pg.REP_KEY64: np.uint64,
              pg.REP_KV16: np.uint32}[self.t.desc.rep]
        return np.frombuffer(self.t.level_bytes(lvl), dtype=dt)
