"""Oracle: the matrix-free operator family ``K_g V = Z_g (Z_g^T V) / eff_g``.

Test infrastructure only.  FP64 brute force on explicit standardized Z;
cites the reference driver each function mirrors.
"""
from __future__ import annotations

import numpy as np

from .geno import OraclePanel, unpack_codes


def _comp_slices(panel: OraclePanel):
    off = panel.component_offsets
    return [slice(int(off[g]), int(off[g + 1])) for g in range(len(panel.component_sizes))]


def kv(panel: OraclePanel, V, normalize=True):
    """kv_impl.py:268-326 (`kv_impl_streamed`): sum over all calls, /eff."""
    Z = panel.standardized()
    V = np.asarray(V, dtype=np.float64)
    out = Z @ (Z.T @ V)
    if normalize and panel.atoms.eff > 0:
        out = out / panel.atoms.eff
    return out


def xtv(panel: OraclePanel, V, normalize=False):
    """kv_impl.py:382-439 (`xtv_impl_streamed`): (m, k) = Z^T V [/eff]."""
    Z = panel.standardized()
    out = Z.T @ np.asarray(V, dtype=np.float64)
    if normalize and panel.atoms.eff > 0:
        out = out / panel.atoms.eff
    return out


def component_kv(panel: OraclePanel, V, g, normalize=True):
    """kv_impl.py:1042-1083 (`kv_impl_partitioned_component`)."""
    Z = panel.standardized()[:, _comp_slices(panel)[g]]
    out = Z @ (Z.T @ np.asarray(V, dtype=np.float64))
    eff = float(panel.atoms.component_eff[g])
    if normalize and eff > 0:
        out = out / eff
    return out


def stacked_kv(panel: OraclePanel, V, normalize=True):
    """kv_impl.py:1136-1187 (`kv_impl_partitioned_stacked`): (G, n, k)."""
    return np.stack([component_kv(panel, V, g, normalize) for g in range(len(panel.component_sizes))])


def weighted_hv(panel: OraclePanel, theta_g, theta_e, V):
    """kv_impl.py:1086-1133 (`kv_impl_partitioned_weighted`):
    ``theta_e V + sum_g (theta_g/eff_g) Z_g Z_g^T V`` (eff_g == 0 -> scale 0)."""
    V = np.asarray(V, dtype=np.float64)
    acc = np.zeros_like(V) if theta_e is None else float(theta_e) * V
    for g in range(len(panel.component_sizes)):
        eff = float(panel.atoms.component_eff[g])
        scale = float(theta_g[g]) / eff if eff > 0 else 0.0
        acc = acc + scale * component_kv(panel, V, g, normalize=False)
    return acc


def core_atoms(panel: OraclePanel, U, subtract_identity=True):
    """kv_impl.py:1287-1339 (`build_projected_core_atoms_partitioned`):
    ``(X_c^T U)^T (X_c^T U)`` summed per component, /eff_g, minus I."""
    U = np.asarray(U, dtype=np.float64)
    Z = panel.standardized()
    r = U.shape[1]
    out = []
    for g, sl in enumerate(_comp_slices(panel)):
        B = Z[:, sl].T @ U
        core = B.T @ B
        eff = float(panel.atoms.component_eff[g])
        if eff > 0:
            core = core / eff
            if subtract_identity:
                core = core - np.eye(r)
        else:
            core = np.zeros_like(core)
        out.append(core)
    return np.stack(out)


def kv_packed_calls(packed_calls, widths, mean, inv_sd, V, eff, normalize=True,
                    dtype=np.float32):
    """Reference-structured CPU K.V over packed calls (kv_impl.py:58-113, 268-326).

    For every call: unpack the (n, pw) bytes, impute missing to the mean and
    centre (kv_impl.py:77-91), ``inner = diff^T V`` and
    ``acc += diff @ (inv^2 * inner)`` in ``dtype`` BLAS GEMMs; finally /eff.
    Used as the timed CPU baseline (FP32, the reference's arithmetic width).
    """
    V = np.asarray(V, dtype=dtype)
    acc = np.zeros_like(V)
    s0 = 0
    for blk, w in zip(packed_calls, widths):
        g = unpack_codes(blk, int(w))
        mu = np.asarray(mean[s0 : s0 + w], dtype=dtype)[None, :]
        iv = np.asarray(inv_sd[s0 : s0 + w], dtype=dtype)
        diff = np.where(g == 3, mu, g.astype(dtype)) - mu
        inner = diff.T @ V
        acc += diff @ ((iv * iv)[:, None] * inner)
        s0 += int(w)
    if normalize and eff > 0:
        acc /= dtype(eff)
    return acc


def exact_diag(panel: OraclePanel):
    """geno_stream.py:2791-2835 (`exact_diag`): per-sample sum_j z_ij^2 / eff."""
    Z = panel.standardized()
    eff = panel.atoms.eff
    return (Z * Z).sum(axis=1) / eff if eff > 0 else np.zeros(panel.n)
