"""Test helpers: host <-> device conversions in the reference's column-major layouts
and the error metrics used by the parity tests."""
import sys

import numpy as np
import torch

import aok


def log(msg):
    """Name the case before a GPU step (flushed, so a fault still leaves a trace)."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def dev_f(a, dtype, device):
    """numpy (logical shape) -> column-major device tensor."""
    return aok.fortran_from_numpy(np.asarray(a, dtype=np.float64), dtype, device)


def dev_rows(a, dtype, device):
    """(n, 3) positions -> contiguous device tensor."""
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), dtype=dtype, device=device)


def rounded(a, dtype):
    """The value the device actually sees (FP32 inputs are rounded)."""
    a = np.asarray(a, dtype=np.float64)
    return a.astype(np.float32).astype(np.float64) if dtype == torch.float32 else a


def host(t):
    return t.detach().cpu().double().numpy()


def rel_max(got, ref):
    """max |got - ref| / max |ref|."""
    d = np.max(np.abs(got - ref)) if got.size else 0.0
    return d / max(np.max(np.abs(ref)) if ref.size else 0.0, 1e-300)


def blockwise_rel(got, ref):
    """max_b max_{ij} |dS| / max_{ij} |S64_b| for (N, N, B) blocks."""
    num = np.max(np.abs(got - ref), axis=(0, 1))
    den = np.maximum(np.max(np.abs(ref), axis=(0, 1)), 1e-300)
    return float(np.max(num / den))


def elementwise_rel(got, ref, tau=1e-3):
    """Relative error over entries with |ref| >= tau * max|ref_b|; returns (max, excluded count)."""
    big = np.max(np.abs(ref), axis=(0, 1), keepdims=True)
    mask = np.abs(ref) >= tau * big
    r = np.abs(got - ref)[mask] / np.abs(ref)[mask]
    return float(r.max() if r.size else 0.0), int(mask.size - mask.sum())
