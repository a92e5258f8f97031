"""Synthetic inputs of the benchmarked workloads, generated on the GPU with
torch (device buffers only): Zipf-distributed sparse tensors with CP-model
values, and ill-conditioned factor matrices.  Shared by bench.py and the GPU
tests so that both run the same data."""
from __future__ import annotations


def zipf_coords(torch, n, size, s, gen, dev):
    """``size`` truncated Zipf(s) indices in [0, n): inverse transform of the
    continuous density x^-s on [1, n+1), floored and shifted to 0."""
    u = torch.rand(size, generator=gen, device=dev, dtype=torch.float64)
    a = 1.0 - (n + 1.0) ** (1.0 - s)
    x = (1.0 - u * a) ** (1.0 / (1.0 - s))
    return (torch.floor(x).to(torch.int64) - 1).clamp_(0, n - 1)


def zipf_sparse(torch, dims, nnz, R, dev, seed=2024, exponent=1.1, noise=0.01, chunk=10_000_000):
    """A sparse tensor of ``len(dims)`` modes on the device: ``nnz`` distinct
    coordinates, each mode's index drawn from a truncated Zipf(``exponent``)
    law (int32, nnz x N), values = the rank-R CP model of N(0,1) factors at
    those coordinates plus ``noise`` x std Gaussian noise (FP64).  The
    coordinates are packed into one 63-bit key per nonzero for the
    deduplication.  Returns (idx, vals)."""
    N = len(dims)
    bits = [(d - 1).bit_length() for d in dims]
    if sum(bits) > 63:
        raise ValueError("the coordinates of one nonzero must pack into 63 bits")
    shift = [sum(bits[k + 1:]) for k in range(N)]
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    keys = torch.empty(0, dtype=torch.int64, device=dev)
    while keys.numel() < nnz:
        m = max(int(1.25 * (nnz - keys.numel())), 1_000_000)
        k = zipf_coords(torch, dims[0], m, exponent, gen, dev) << shift[0]
        for q in range(1, N):
            k |= zipf_coords(torch, dims[q], m, exponent, gen, dev) << shift[q]
        keys = torch.unique(torch.cat([keys, k]))
        del k
    keys = keys[torch.randperm(keys.numel(), generator=gen, device=dev)[:nnz]]
    idx = torch.stack([(keys >> shift[q]) & ((1 << bits[q]) - 1) for q in range(N)], dim=1)
    del keys
    idx = idx.to(torch.int32).contiguous()
    A = [torch.randn((d, R), generator=gen, device=dev, dtype=torch.float64) for d in dims]
    vals = torch.empty(nnz, dtype=torch.float64, device=dev)
    for s0 in range(0, nnz, chunk):
        s1 = min(nnz, s0 + chunk)
        p = A[0][idx[s0:s1, 0].long()]
        for k in range(1, N):
            p *= A[k][idx[s0:s1, k].long()]
        vals[s0:s1] = p.sum(dim=1)
    vals += noise * vals.std() * torch.randn(nnz, generator=gen, device=dev, dtype=torch.float64)
    del A
    return idx, vals


def ill_conditioned_factors(torch, N, I, R, dev, seed, smin=1e-6):
    """N factors U = Q_U diag(s) Q_V^T (FP64, I x R, on the device): Q_U (I x R)
    and Q_V (R x R) orthonormal from the QR of seeded Gaussian matrices, s
    geometric from 1 to ``smin``."""
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    s = torch.logspace(0, torch.log10(torch.tensor(smin)).item(), R, dtype=torch.float64, device=dev)
    out = []
    for _ in range(N):
        qu, _ = torch.linalg.qr(torch.randn((I, R), generator=gen, device=dev, dtype=torch.float64))
        qv, _ = torch.linalg.qr(torch.randn((R, R), generator=gen, device=dev, dtype=torch.float64))
        out.append(((qu * s) @ qv.T).contiguous())
        del qu, qv
    return out, gen
