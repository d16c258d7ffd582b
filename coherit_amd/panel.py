"""Resident genotype panel: a thin owner of the C-ABI ``coh_panel`` handle.

torch supplies device memory and the current stream; every byte of genotype
decoding and contraction happens in the HIP library.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _native as N


# When set to a list, every operator call appends (op, k, n, m_sel, precision):
# the genotype passes of a workload, e.g. to extrapolate a CPU baseline.
CALL_LOG: list | None = None


def _log(op, k, n, m_sel, precision):
    if CALL_LOG is not None:
        CALL_LOG.append((op, int(k), int(n), int(m_sel), int(precision)))


def _dev(device) -> torch.device:
    if device is None or device == "auto":
        return torch.device("cuda", torch.cuda.current_device())
    d = torch.device(device)
    if d.type != "cuda":
        raise ValueError(f"coherit_amd panels live on a ROCm device, got {device!r}")
    return d if d.index is not None else torch.device("cuda", torch.cuda.current_device())


class ResidentPanel:
    """Packed panel resident in HBM in sample-major and variant-major layouts."""

    def __init__(self, n: int, call_starts, call_widths, call_components, n_components: int,
                 device=None):
        N.lib()
        self.device = _dev(device)
        starts = np.ascontiguousarray(call_starts, dtype=np.int64)
        widths = np.ascontiguousarray(call_widths, dtype=np.int32)
        comps = np.ascontiguousarray(call_components, dtype=np.int32)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            N.call("coh_panel_create", int(n), int(starts.size), N.ptr(starts), N.ptr(widths),
                   N.ptr(comps), int(n_components), ctypes.byref(handle))
        self._h = handle
        info = np.zeros(8, np.int64)
        N.call("coh_panel_info", self._h, N.ptr(info))
        self.n, self.m, self.n_calls, self.n_components = (int(x) for x in info[:4])
        self.m_slot, self.sbytes, self.vbytes = (int(x) for x in info[4:7])
        b = np.zeros(self.n_components, np.int64)
        e = np.zeros(self.n_components, np.int64)
        N.call("coh_panel_component_slots", self._h, N.ptr(b), N.ptr(e))
        self.comp_slot_begin, self.comp_slot_end = b, e
        cs = np.zeros(self.n_calls, np.int64)
        N.call("coh_panel_call_slots", self._h, N.ptr(cs))
        self.call_slot = cs
        self.call_starts = starts
        self.call_widths = widths
        self.call_components = comps
        self.finalized = False

    # ------------------------------------------------------------ lifetime
    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("GenoBlockStreamer is closed.")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            torch.cuda.synchronize(self.device)
            N.lib().coh_panel_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def nbytes(self) -> int:
        return self.n * self.sbytes + self.m_slot * self.vbytes

    # -------------------------------------------------------------- ingest
    def ingest_packed_call(self, c: int, block: torch.Tensor):
        assert block.dtype == torch.uint8 and block.is_cuda and block.dim() == 2
        N.call("coh_panel_ingest_packed_call", self.handle, int(c), N.ptr(block),
               int(block.stride(0)), N.stream_ptr())

    def ingest_i8_rows(self, start: int, block: torch.Tensor, missing_raw: int):
        assert block.dtype == torch.int8 and block.is_cuda and block.dim() == 2
        assert block.stride(1) == 1
        N.call("coh_panel_ingest_i8_varmaj", self.handle, int(start), int(block.shape[0]),
               N.ptr(block), int(block.stride(0)), int(missing_raw), N.stream_ptr())

    def ingest_bed_rows(self, start: int, rows: torch.Tensor, bytes_per_snp: int,
                        sample_idx: torch.Tensor):
        assert rows.dtype == torch.uint8 and rows.is_cuda and rows.is_contiguous()
        assert sample_idx.dtype == torch.int32 and sample_idx.numel() == self.n
        N.call("coh_panel_ingest_bed", self.handle, int(start), int(rows.shape[0]), N.ptr(rows),
               int(bytes_per_snp), N.ptr(sample_idx), N.stream_ptr())

    def generate(self, seed: int, snp_ids, t0, t1, tmiss: int):
        d = self.device
        ids = torch.as_tensor(np.asarray(snp_ids, np.uint32).view(np.int32), device=d)
        a = torch.as_tensor(np.asarray(t0, np.uint32).view(np.int32), device=d)
        b = torch.as_tensor(np.asarray(t1, np.uint32).view(np.int32), device=d)
        N.call("coh_panel_generate", self.handle, int(seed) & 0xFFFFFFFFFFFFFFFF, N.ptr(ids),
               N.ptr(a), N.ptr(b), int(tmiss), N.stream_ptr())
        torch.cuda.current_stream(d).synchronize()

    def finalize(self, mean_override=None, inv_override=None):
        mo = io = None
        if mean_override is not None:
            mo = torch.as_tensor(np.asarray(mean_override, np.float32), device=self.device)
            io = torch.as_tensor(np.asarray(inv_override, np.float32), device=self.device)
        N.call("coh_panel_finalize", self.handle, N.ptr(mo), N.ptr(io), N.stream_ptr())
        torch.cuda.current_stream(self.device).synchronize()
        self.finalized = True

    def stats(self):
        """Cache-order (mean, inv_sd, cnt, s1, s2) as host NumPy arrays."""
        d = self.device
        mean = torch.empty(self.m, dtype=torch.float32, device=d)
        inv = torch.empty_like(mean)
        cnt = torch.empty(self.m, dtype=torch.int64, device=d)
        s1 = torch.empty_like(cnt)
        s2 = torch.empty_like(cnt)
        N.call("coh_panel_copy_stats", self.handle, N.ptr(mean), N.ptr(inv), N.ptr(cnt), N.ptr(s1),
               N.ptr(s2), N.stream_ptr())
        return tuple(t.cpu().numpy() for t in (mean, inv, cnt, s1, s2))

    @property
    def cache_to_slot_host(self) -> np.ndarray:
        """(m,) int64 slot of every cache-order SNP."""
        if getattr(self, "_c2s", None) is None:
            c2s = np.empty(self.m, np.int64)
            for c in range(self.n_calls):
                a, w = int(self.call_starts[c]), int(self.call_widths[c])
                c2s[a:a + w] = int(self.call_slot[c]) + np.arange(w)
            self._c2s = c2s
        return self._c2s

    def set_standardization(self, mean, inv_sd) -> None:
        """Replace mean / inv_sd (cache order, m entries) of the finalized panel."""
        def dev(x):
            if isinstance(x, torch.Tensor):
                return x.to(device=self.device, dtype=torch.float32).reshape(-1).contiguous()
            return torch.as_tensor(np.asarray(x, np.float32).reshape(-1), device=self.device)
        mo, io = dev(mean), dev(inv_sd)
        if mo.numel() != self.m or io.numel() != self.m:
            raise ValueError(f"standardization needs {self.m} entries per array")
        N.call("coh_panel_set_standardization", self.handle, N.ptr(mo), N.ptr(io), N.stream_ptr())
        torch.cuda.current_stream(self.device).synchronize()

    def read_packed_call(self, c: int) -> torch.Tensor:
        pw = (int(self.call_widths[c]) + 3) // 4
        out = torch.empty((self.n, pw), dtype=torch.uint8, device=self.device)
        N.call("coh_panel_read_packed_call", self.handle, int(c), N.ptr(out), pw, N.stream_ptr())
        return out

    # ----------------------------------------------------------- operators
    def _m_sel(self, cb, ce):
        sel = (self.call_components >= cb) & (self.call_components < ce)
        return int(self.call_widths[sel].astype(np.int64).sum())

    def workspace_bytes(self, op: int, precision: int, k: int, cb: int, ce: int) -> int:
        nbytes = ctypes.c_uint64()
        N.call("coh_workspace_bytes", self.handle, int(op), int(precision), int(k), int(cb),
               int(ce), ctypes.byref(nbytes))
        return int(nbytes.value)

    def _workspace(self, op, precision, k, cb, ce):
        nb = self.workspace_bytes(op, precision, k, cb, ce)
        return torch.empty(max(nb, 256), dtype=torch.uint8, device=self.device), nb

    def kv(self, V: torch.Tensor, comp_scale: Sequence[float], cb: int, ce: int, *,
           stacked: bool = False, out: torch.Tensor | None = None, beta: float = 0.0,
           precision: int = N.PREC_FP32) -> torch.Tensor:
        """(beta*out) + sum_g comp_scale[g] Z_g Z_g^T V over components [cb, ce)."""
        V = _as_f32_matrix(V, self.device, self.n)
        k = int(V.shape[1])
        G = ce - cb
        if out is None:
            shape = (G, self.n, k) if stacked else (self.n, k)
            out = torch.empty(shape, dtype=torch.float32, device=self.device)
            beta = 0.0
        scale = np.ascontiguousarray(comp_scale, dtype=np.float32).reshape(-1)
        if scale.size != self.n_components:
            raise ValueError(f"comp_scale needs {self.n_components} entries, got {scale.size}")
        ws, nb = self._workspace(N.OP_KV, precision, k, cb, ce)
        if stacked:
            ldo, ostride = int(out.stride(1)), int(out.stride(0))
        else:
            ldo, ostride = int(out.stride(0)), 0
        _log("kv", k, self.n, self._m_sel(cb, ce), precision)
        N.call("coh_kv", self.handle, int(precision), N.ptr(V), k, int(V.stride(0)), int(cb), int(ce),
               N.ptr(scale), int(bool(stacked)), N.ptr(out), ldo, ostride, float(beta), N.ptr(ws),
               nb, N.stream_ptr())
        return out

    def xtv(self, V: torch.Tensor, cb: int, ce: int, *, scale: float = 1.0,
            slot_space: bool = False, precision: int = N.PREC_FP32) -> torch.Tensor:
        V = _as_f32_matrix(V, self.device, self.n)
        k = int(V.shape[1])
        if slot_space:
            rows = int(self.comp_slot_end[ce - 1] - self.comp_slot_begin[cb])
        else:
            sel = (self.call_components >= cb) & (self.call_components < ce)
            rows = int(self.call_widths[sel].astype(np.int64).sum())
        out = torch.empty((rows, k), dtype=torch.float32, device=self.device)
        ws, nb = self._workspace(N.OP_XTV, precision, k, cb, ce)
        _log("xtv", k, self.n, self._m_sel(cb, ce), precision)
        N.call("coh_xtv", self.handle, int(precision), N.ptr(V), k, int(V.stride(0)), int(cb),
               int(ce), float(scale), int(bool(slot_space)), N.ptr(out), int(out.stride(0)),
               N.ptr(ws), nb, N.stream_ptr())
        return out

    def zb(self, B: torch.Tensor, cb: int, ce: int, *, stacked: bool = False,
           out: torch.Tensor | None = None, beta: float = 0.0,
           precision: int = N.PREC_FP32) -> torch.Tensor:
        """(beta*out) + Z_g B_g with B rows = cache SNPs of components [cb, ce)."""
        if not isinstance(B, torch.Tensor):
            B = torch.as_tensor(np.asarray(B))
        B = B.to(device=self.device, dtype=torch.float32).contiguous()
        if B.dim() != 2:
            raise ValueError("B must be 2-D (m_sel, k)")
        sel = (self.call_components >= cb) & (self.call_components < ce)
        m_sel = int(self.call_widths[sel].astype(np.int64).sum())
        if int(B.shape[0]) != m_sel:
            raise ValueError(f"B has {int(B.shape[0])} rows; components [{cb}, {ce}) hold {m_sel} SNPs")
        k = int(B.shape[1])
        G = ce - cb
        if out is None:
            out = torch.empty((G, self.n, k) if stacked else (self.n, k), dtype=torch.float32,
                              device=self.device)
            beta = 0.0
        ws, nb = self._workspace(N.OP_KV, precision, k, cb, ce)
        ldo, ostride = (int(out.stride(1)), int(out.stride(0))) if stacked else (int(out.stride(0)), 0)
        _log("zb", k, self.n, m_sel, precision)
        N.call("coh_zb", self.handle, int(precision), N.ptr(B), k, int(B.stride(0)), int(cb), int(ce),
               int(bool(stacked)), N.ptr(out), ldo, ostride, float(beta), N.ptr(ws), nb,
               N.stream_ptr())
        return out

    def exact_diag(self, eff: float) -> torch.Tensor:
        out = torch.empty(self.n, dtype=torch.float32, device=self.device)
        N.call("coh_exact_diag", self.handle, float(eff), N.ptr(out), N.stream_ptr())
        return out


def _as_f32_matrix(V: torch.Tensor, device, n: int) -> torch.Tensor:
    if not isinstance(V, torch.Tensor):
        V = torch.as_tensor(np.asarray(V))
    if V.dim() != 2:
        raise ValueError(f"expected a 2-D (n, k) block, got shape {tuple(V.shape)}")
    if int(V.shape[0]) != n:
        raise ValueError(f"V has {int(V.shape[0])} rows; the panel has n={n} samples.")
    if V.device != device or V.dtype != torch.float32 or V.stride(1) != 1:
        V = V.to(device=device, dtype=torch.float32).contiguous()
    return V
