"""Benchmark of the batched Cholesky solve (potrf + two trsm) at n = 128.

Workload: 8192 independent SPD systems A = G G^T + 128 I (G ~ N(0,1)),
4 right-hand sides U(-1,1) each, BLAST panel storage (PS = 4).  One step is
one batched call over the whole batch with every input already in HBM:

  plain FP64  blast_hip_dposv_batched: A (lower) in, L (lower) and X out,
              the LAPACK dposv contract;
  mixed + IR  blast_hip_dsposv_batched: factor in FP32 (or BF16 products),
              refine in FP64, X out (LAPACK dsposv contract).

`value` is the rate of the faster of the two paths among those whose every
solve meets the 1e-12 backward-error bar (config.headline_path names it);
both paths, their errors and rooflines are reported.  The line also
carries the other workloads (dgemm 64^3, storage conversion, the uneven
Riccati recursion, n = 256 mixed precision, adjoints) under
config.other_workloads.

Prints one JSON line.  Time is measured with HIP events on the stream the
kernels run on, over --steps steps after --warmup steps.  The CPU baseline is
the oracle (single-threaded BLAST restatement), one instance per host core,
timed on a bounded sample of the same batch.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, NRHS, BATCH = 128, 4, 8192
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
# Algorithmic work per instance (potrf n^3/3 + n^2/2 + n/6; each trsm n^2 * nrhs).
FLOPS_PER_INSTANCE = N ** 3 / 3 + N ** 2 / 2 + N / 6 + 2 * N * N * NRHS
# Lower triangle of A in, lower triangle of L out, B in, X out (8-byte words).
TRI = N * (N + 1) // 2
BYTES_PLAIN = 8 * (2 * TRI + 2 * N * NRHS)
BYTES_SOLVE_ONLY = 8 * (TRI + 2 * N * NRHS)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--batch", type=int, default=BATCH)
    p.add_argument("--cpu-seconds", type=float, default=10.0, help="target length of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--riccati-count", type=int, default=20000, help="instances of the uneven Riccati workload")
    p.add_argument("--only", choices=("all", "posv", "gemm", "riccati", "mixed256", "adjoint", "convert", "blocked",
                                      "blocked_lu"),
                   default="all",
                   help="run one workload alone (diagnostics; the bench line needs 'all')")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="first run a 256 MiB device copy (known HBM bytes) to scale rocprofv3 FETCH/WRITE_SIZE")
    return p.parse_args()


def make_inputs(batch, seed=20240611):
    """SPD batch and RHS generated on the GPU (torch Philox, fixed seed), packed
    into panel storage on the device."""
    import torch

    from blast_amd.layout import Layout

    gen = torch.Generator(device="cuda").manual_seed(seed)
    la, lb = Layout.make(N, N), Layout.make(N, NRHS)
    abuf = torch.zeros(batch * la.batch_stride, dtype=torch.float64, device="cuda")
    bbuf = torch.zeros(batch * lb.batch_stride, dtype=torch.float64, device="cuda")
    ia = torch.from_numpy(la.index_map().reshape(-1)).cuda()
    ib = torch.from_numpy(lb.index_map().reshape(-1)).cuda()
    eye = N * torch.eye(N, dtype=torch.float64, device="cuda")
    chunk = 1024
    for s in range(0, batch, chunk):
        e = min(batch, s + chunk)
        g = torch.randn((e - s, N, N), dtype=torch.float64, device="cuda", generator=gen)
        a = torch.bmm(g, g.transpose(1, 2).contiguous()) + eye
        abuf.view(batch, la.batch_stride)[s:e, :].index_copy_(1, ia, a.reshape(e - s, -1))
    b = torch.rand((batch, N, NRHS), dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    bbuf.view(batch, lb.batch_stride)[:, :].index_copy_(1, ib, b.reshape(batch, -1))
    return la, lb, abuf, bbuf


def backward_errors(la, lb, abuf, xbuf, bbuf, batch, lbuf=None):
    """Per-instance normwise backward errors, computed in FP64 on the GPU with
    torch (checker only)."""
    import torch

    ia = torch.from_numpy(la.index_map()).cuda()
    ib = torch.from_numpy(lb.index_map()).cuda()
    out_solve, out_fact = [], []
    for s in range(0, batch, 1024):
        e = min(batch, s + 1024)
        a = abuf.view(batch, la.batch_stride)[s:e][:, ia]
        a = torch.tril(a) + torch.tril(a, -1).transpose(1, 2)
        x = xbuf.view(batch, lb.batch_stride)[s:e][:, ib]
        b = bbuf.view(batch, lb.batch_stride)[s:e][:, ib]
        r = b - torch.bmm(a, x)
        na, nx, nb = (torch.linalg.matrix_norm(t) for t in (a, x, b))
        out_solve.append((torch.linalg.matrix_norm(r) / (na * nx + nb)).cpu())
        if lbuf is not None:
            l = torch.tril(lbuf.view(batch, la.batch_stride)[s:e][:, ia])
            out_fact.append((torch.linalg.matrix_norm(torch.bmm(l, l.transpose(1, 2).contiguous()) - a) / na).cpu())
    se = torch.cat(out_solve).numpy()
    fe = torch.cat(out_fact).numpy() if out_fact else None
    return se, fe


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    import re

    path = os.path.join(ROOT, "profiles", "posv_n128_hbm_traffic.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch " + re.escape(kernel) + r":\s*(\d+)", text)
    return int(m.group(1)) if m else None


def time_steps(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / 1e3 / steps  # seconds per step


def gemm_workload(steps, warmup, batch=4096, n=64):
    """Batched dgemm NN D = 1.5 A B - 0.5 C, m = n = k = 64, every operand in
    StaticPanelMatrix layout, U(-1,1) inputs (the BLAST panel gemm benchmark
    shape at n = 64)."""
    import torch

    from blast_amd import capi
    from blast_amd.layout import Layout

    gen = torch.Generator(device="cuda").manual_seed(7)
    la = Layout.make(n, n)
    bufs = [torch.rand(batch * la.batch_stride, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
            for _ in range(4)]
    A, B, C, D = (capi.Operand(la, t) for t in bufs)

    def step():
        capi.gemm(1.5, A, B, -0.5, C, D, batch)

    t = time_steps(step, steps, warmup)
    flops = 2 * n ** 3 * batch
    hbm = 4 * n * n * 8 * batch
    return {"gflops": flops / t / 1e9, "instances_per_s": batch / t, "ms_per_step": t * 1e3,
            "hbm_gbs_algorithmic": hbm / t / 1e9, "hbm_frac": hbm / t / 1e9 / PEAK_HBM_GBS,
            "kernel": "gemm_small_kernel<double>"}


def blocked_workload(steps, warmup, batch=256, n=512, nrhs=4):
    """dposv and dgesv beyond the n <= 256 kernels: n = 512, 4 RHS, batch
    256.  dposv: panel storage, A = G G^T + n I, the AUTO blocked path
    (256-row blocks on the MFMA kernels); dgesv: the blocked LU.  Each
    against the reference-order column sweeps it replaces (timed once: they
    are slow), with the backward error of both."""
    import torch

    from blast_amd import capi
    from blast_amd.layout import Layout

    gen = torch.Generator(device="cuda").manual_seed(11)
    g = torch.randn(batch, n, n, dtype=torch.float64, device="cuda", generator=gen)
    a = torch.bmm(g, g.transpose(1, 2).contiguous()) + n * torch.eye(n, dtype=torch.float64, device="cuda")
    b = torch.rand(batch, n, nrhs, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    la, lb = Layout.make(n, n), Layout.make(n, nrhs)

    def pack(x, lay):
        idx = torch.from_numpy(lay.index_map().reshape(-1)).cuda()
        buf = torch.zeros(batch * lay.batch_stride, dtype=torch.float64, device="cuda")
        buf.view(batch, lay.batch_stride).index_copy_(1, idx, x.reshape(batch, -1))
        return capi.Operand(lay, buf)

    A, B = pack(a, la), pack(b, lb)
    L, X = capi.Operand(la, torch.zeros_like(A.tensor)), capi.Operand(lb, torch.zeros_like(B.tensor))
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    ib = torch.from_numpy(lb.index_map().reshape(-1)).cuda()

    def bwd():
        x = X.tensor.view(batch, lb.batch_stride)[:, ib].reshape(batch, n, nrhs)
        r = b - torch.bmm(a, x)
        return float((torch.linalg.matrix_norm(r) / (torch.linalg.matrix_norm(a) * torch.linalg.matrix_norm(x)
                                                   + torch.linalg.matrix_norm(b))).max())

    out = {"batch": batch, "n": n, "nrhs": nrhs}
    capi.set_algorithm(capi.ALGO_AUTO)
    t = time_steps(lambda: capi.posv(A, L, B, X, batch, info), steps, warmup)
    flops = batch * (n ** 3 / 3 + 2 * n * n * nrhs)
    out["auto_blocked"] = {"solves_per_s": batch / t, "ms_per_step": t * 1e3, "fp64_tflops": flops / t / 1e12,
                           "max_solve_backward_error": bwd(), "info_nonzero": int((info != 0).sum())}
    capi.set_algorithm(capi.ALGO_REFERENCE)
    try:
        t_ref = time_steps(lambda: capi.posv(A, L, B, X, batch, info), 1, 0)
    finally:
        capi.set_algorithm(capi.ALGO_AUTO)
    out["reference_column_sweeps"] = {"solves_per_s": batch / t_ref, "ms_per_step": t_ref * 1e3,
                                      "max_solve_backward_error": bwd()}

    out["dgesv"] = blocked_lu_workload(steps, warmup, batch, n, nrhs)
    return out


def blocked_lu_workload(steps, warmup, batch=256, n=512, nrhs=4):
    """dgesv (LU with partial pivoting, in place of A) beyond the n <= 256
    kernel: n = 512, 4 RHS, batch 256, dense column-major, A ~ U(-1, 1) +
    I/2: the AUTO blocked LU (register panels of 24 columns + gemm tile
    updates) + getrs against the reference column sweeps (timed once).
    Each step first restores A from a pristine copy (timed alone as
    copy_ms and subtracted)."""
    import torch

    from blast_amd import capi
    from blast_amd.layout import Layout

    gen = torch.Generator(device="cuda").manual_seed(12)
    b = torch.rand(batch, n, nrhs, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")

    def pack(x, lay):
        idx = torch.from_numpy(lay.index_map().reshape(-1)).cuda()
        buf = torch.zeros(batch * lay.batch_stride, dtype=torch.float64, device="cuda")
        buf.view(batch, lay.batch_stride).index_copy_(1, idx, x.reshape(batch, -1))
        return capi.Operand(lay, buf)

    lc, lbc = Layout.make(n, n, panel=False), Layout.make(n, nrhs, panel=False)
    ag = torch.rand(batch, n, n, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    ag += 0.5 * torch.eye(n, dtype=torch.float64, device="cuda")
    A0, Bg = pack(ag, lc), pack(b, lbc)
    Aw, Xg = capi.Operand(lc, torch.empty_like(A0.tensor)), capi.Operand(lbc, torch.zeros_like(Bg.tensor))
    ipiv = torch.zeros(batch * n, dtype=torch.int64, device="cuda")
    ibc = torch.from_numpy(lbc.index_map().reshape(-1)).cuda()

    def gesv_step():
        Aw.tensor.copy_(A0.tensor)
        capi.gesv(Aw, ipiv, n, Bg, Xg, batch, info)

    def bwd_g():
        x = Xg.tensor.view(batch, lbc.batch_stride)[:, ibc].reshape(batch, n, nrhs)
        r = b - torch.bmm(ag, x)
        return float((torch.linalg.matrix_norm(r) / (torch.linalg.matrix_norm(ag) * torch.linalg.matrix_norm(x)
                                                   + torch.linalg.matrix_norm(b))).max())

    t_copy = time_steps(lambda: Aw.tensor.copy_(A0.tensor), steps, warmup)
    t = time_steps(gesv_step, steps, warmup)
    gflops = batch * (2 * n ** 3 / 3 + 2 * n * n * nrhs)
    g = {"layout": "dense column-major", "copy_ms": t_copy * 1e3, "ms_per_step_incl_copy": t * 1e3,
         "ms_per_step": (t - t_copy) * 1e3, "solves_per_s": batch / (t - t_copy),
         "fp64_tflops": gflops / (t - t_copy) / 1e12, "max_solve_backward_error": bwd_g(),
         "info_nonzero": int((info != 0).sum())}
    capi.set_algorithm(capi.ALGO_REFERENCE)
    try:
        t_ref = time_steps(gesv_step, 1, 0)
    finally:
        capi.set_algorithm(capi.ALGO_AUTO)
    g["reference_column_sweeps"] = {"ms_per_step_incl_copy": t_ref * 1e3, "solves_per_s": batch / (t_ref - t_copy),
                                    "max_solve_backward_error": bwd_g()}
    return g


def convert_workload(steps, warmup, batch=4096, n=128):
    """Storage conversion of 4096 FP64 128 x 128 matrices: panel -> dense
    column-major, and panel -> dense row-major (transposing); HBM-bound,
    8 * n^2 bytes read + 8 * n^2 written per instance."""
    import torch

    from blast_amd import capi
    from blast_amd.layout import Layout

    lp, lc, lr = Layout.make(n, n), Layout.make(n, n, panel=False), Layout.make(n, n, panel=False).t()
    gen = torch.Generator(device="cuda").manual_seed(3)
    S = capi.Operand(lp, torch.rand(batch * lp.batch_stride, dtype=torch.float64, device="cuda", generator=gen))
    out = {}
    for name, ld in (("panel_to_colmajor", lc), ("panel_to_rowmajor", lr)):
        D = capi.Operand(ld, torch.zeros(batch * ld.batch_stride, dtype=torch.float64, device="cuda"))
        t = time_steps(lambda: capi.copy(S, D, batch), steps, warmup)
        gbs = 2 * 8 * n * n * batch / t / 1e9
        out[name] = {"ms_per_step": t * 1e3, "hbm_gbs_algorithmic": gbs, "hbm_frac": gbs / PEAK_HBM_GBS}
    out["kernel"] = "convert_kernel<double>"
    return out


def riccati_inputs(count=20000, horizon=10, seed=20240611):
    """Uneven Riccati batch: nx ~ U{4..200}, nu ~ U{1..nx/2}, A, B ~ N(0, 1/nx),
    Q = I + 0.1 W W^T (W ~ N(0,1)), R = I, generated on the GPU per nx group."""
    import torch

    from blast_amd import capi

    rng = np.random.default_rng(seed)
    nx = rng.integers(4, 201, size=count).astype(np.int32)
    nu = rng.integers(1, np.maximum(1, nx // 2) + 1).astype(np.int32)
    batch = capi.RiccatiBatch(nx, nu, horizon)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    inst = batch.inst
    buf = batch.input
    for n in np.unique(nx):
        sel = np.nonzero(nx == n)[0]
        g, n = len(sel), int(n)
        m_max = max(1, n // 2)
        m = torch.from_numpy(nu[sel].astype(np.int64)).cuda()
        cols_nn = torch.arange(n * n, device="cuda")
        off = {k: torch.from_numpy(inst[k][sel]).cuda() for k in ("a", "b", "q", "r")}
        a = torch.randn(g, n, n, dtype=torch.float64, device="cuda", generator=gen) / np.sqrt(n)
        buf[(off["a"][:, None] + cols_nn[None]).reshape(-1)] = a.transpose(1, 2).reshape(-1)
        w = torch.randn(g, n, n, dtype=torch.float64, device="cuda", generator=gen)
        q = torch.eye(n, dtype=torch.float64, device="cuda") + 0.1 * torch.bmm(w, w.transpose(1, 2).contiguous())
        buf[(off["q"][:, None] + cols_nn[None]).reshape(-1)] = q.transpose(1, 2).reshape(-1)
        bm = torch.randn(g, m_max, n, dtype=torch.float64, device="cuda", generator=gen) / np.sqrt(n)
        flat = torch.arange(m_max * n, device="cuda")
        keep = (flat[None] // n) < m[:, None]  # column j = flat // n of B kept when j < nu
        idx = off["b"][:, None] + flat[None]
        buf[idx[keep]] = bm.reshape(g, -1)[keep]
        diag = torch.arange(m_max, device="cuda")
        keep = diag[None] < m[:, None]
        ridx = off["r"][:, None] + diag[None] * (m[:, None] + 1)
        buf[ridx[keep]] = 1.0
    torch.cuda.synchronize()
    return batch


def riccati_workload(steps, warmup, count=20000, horizon=10):
    """One persistent launch over the whole uneven batch, eager and replayed
    from a captured hipGraph; validated on a sample against the FP64 recursion."""
    import torch

    from blast_amd import capi
    from oracle import riccati

    batch = riccati_inputs(count, horizon)
    nx, nu = batch.inst["nx"].astype(np.float64), batch.inst["nu"].astype(np.float64)
    flops = float(np.sum(horizon * riccati.step_flops(nx, nu) + nx ** 3 / 3))
    io_bytes = 8 * float(np.sum(2 * nx * nx + nx * nu + nu * nu + horizon * nu * nx + nx * nx))
    ms = batch.autotune(reps=2)
    t_eager = time_steps(batch.run, steps, warmup)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        batch.run(stream=s.cuda_stream)
        s.synchronize()
        with torch.cuda.graph(graph, stream=s):
            batch.run(stream=s.cuda_stream)
    torch.cuda.current_stream().wait_stream(s)
    batch.output.zero_()
    t_graph = time_steps(graph.replay, steps, warmup)
    # validate a sample of instances (all sizes represented) against the oracle
    host_in = batch.input.cpu().numpy()
    K, L0 = batch.results()
    info = batch.info.cpu().numpy()
    worst = 0.0
    for i in np.linspace(0, count - 1, 24).astype(int):
        d = batch.inst[i]
        x, u = int(d["nx"]), int(d["nu"])
        get = lambda k, r, c: host_in[d[k]:d[k] + r * c].reshape(c, r).T  # noqa: E731
        Kr, Pr = riccati.riccati(get("a", x, x), get("b", x, u), get("q", x, x), get("r", u, u), horizon)
        worst = max(worst, np.abs(K[i] - Kr).max() / np.abs(Kr).max(),
                    np.abs(L0[i] @ L0[i].T - Pr).max() / np.abs(Pr).max())
    return {
        "instances": count, "horizon": horizon,
        "instances_per_s_graph": count / t_graph, "instances_per_s_eager": count / t_eager,
        "ms_per_step_graph": t_graph * 1e3, "ms_per_step_eager": t_eager * 1e3,
        "gflops": flops / t_graph / 1e9, "fp64_frac_of_peak": flops / t_graph / 1e12 / PEAK_FP64_TFLOPS,
        "io_gbs": io_bytes / t_graph / 1e9,
        "max_rel_err_sample_vs_fp64_recursion": float(worst),
        "info_nonzero": int((info[:count] != 0).sum()),
        "tuning_strip_width": capi.riccati_get_tuning().tolist(),
        "autotune_ms": {str(k): [round(float(v), 4) for v in ms[k]] for k in range(16) if ms[k].any()},
        "kernel": "riccati_persistent_kernel",
    }


def conditioned_inputs(batch, n, nrhs, spd, seed):
    """A = U diag(s) V^T (V = U for the SPD variant), U, V Haar (QR of a
    Gaussian matrix, signs of R folded in), s log-uniform in [1/kappa, 1] with
    both ends present, kappa log-uniform in [1e2, 1e8] per instance; B ~ U(-1,1).
    Generated on the GPU; returns dense column-major A (batch, n, n), B and kappa."""
    import torch

    gen = torch.Generator(device="cuda").manual_seed(seed)
    kappa = 10.0 ** (2.0 + 6.0 * torch.rand(batch, dtype=torch.float64, device="cuda", generator=gen))
    a = torch.empty(batch, n, n, dtype=torch.float64, device="cuda")

    def haar(c):
        q, r = torch.linalg.qr(torch.randn(c, n, n, dtype=torch.float64, device="cuda", generator=gen))
        return q * torch.sign(torch.diagonal(r, dim1=-2, dim2=-1))[:, None, :]

    for s0 in range(0, batch, 256):
        c = min(256, batch - s0)
        u = haar(c)
        v = u if spd else haar(c)
        t = torch.rand(c, n, dtype=torch.float64, device="cuda", generator=gen)
        t[:, 0], t[:, -1] = 0.0, 1.0
        sv = torch.exp(-t * torch.log(kappa[s0:s0 + c])[:, None])
        m = torch.bmm(u * sv[:, None, :], v.transpose(1, 2).contiguous())
        a[s0:s0 + c] = 0.5 * (m + m.transpose(1, 2)) if spd else m
    b = torch.rand(batch, n, nrhs, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    return a, b, kappa


def mixed256_workload(steps, warmup, batch=2048, n=256, nrhs=8):
    """Mixed precision factor + FP64 refinement at n = 256, 8 RHS, cond 1e2..1e8:
    SPD (Cholesky: FP64 dposv vs dsposv FP32 / BF16) and general (LU with
    partial pivoting: FP64 dgesv vs dsgesv FP32 / BF16).  Each timed call
    starts from the pristine A (restored outside the timed region: the FP64
    LU and every fallback write their factor into A)."""
    import torch

    from blast_amd import capi
    from blast_amd.layout import Layout
    from oracle import oracle

    out = {"batch": batch, "n": n, "nrhs": nrhs, "kappa": "log-uniform [1e2, 1e8] per instance",
           "stopping_tol": 1e-13}
    lb = Layout.make(n, nrhs)
    ib = torch.from_numpy(lb.index_map().reshape(-1)).cuda()
    work = capi.mixed_work(n, batch)
    for variant, spd in (("spd", True), ("general", False)):
        a, b, kappa = conditioned_inputs(batch, n, nrhs, spd, seed=20240611 + spd)
        la = Layout.make(n, n, panel=spd)
        ia = torch.from_numpy(la.index_map().reshape(-1)).cuda()
        a0 = torch.zeros(batch * la.batch_stride, dtype=torch.float64, device="cuda")
        a0.view(batch, la.batch_stride).index_copy_(1, ia, a.reshape(batch, -1))
        bbuf = torch.zeros(batch * lb.batch_stride, dtype=torch.float64, device="cuda")
        bbuf.view(batch, lb.batch_stride).index_copy_(1, ib, b.reshape(batch, -1))
        abuf, lbuf = torch.empty_like(a0), torch.empty_like(a0)
        xbuf = torch.zeros_like(bbuf)
        A, L, B, X = (capi.Operand(la, abuf), capi.Operand(la, lbuf), capi.Operand(lb, bbuf),
                      capi.Operand(lb, xbuf))
        it = torch.zeros(batch, dtype=torch.int32, device="cuda")
        info = torch.zeros(batch, dtype=torch.int32, device="cuda")
        ipiv = torch.zeros(batch * n, dtype=torch.int64, device="cuda")
        kdec = torch.floor(torch.log10(kappa)).cpu().numpy().astype(int)
        res = {}
        for name, prec in (("fp64", None), ("fp32_ir", 0), ("bf16_ir", 1)):
            def call():
                if prec is None:
                    if spd:
                        capi.posv(A, L, B, X, batch, info)
                    else:
                        capi.gesv(A, ipiv, n, B, X, batch, info)
                elif spd:
                    capi.sposv_ws(prec, A, B, X, work, batch, it, info, tol=1e-13)
                else:
                    capi.sgesv(prec, A, ipiv, n, B, X, work, batch, it, info, tol=1e-13)

            total = 0.0
            for k in range(warmup + steps):
                abuf.copy_(a0)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                call()
                e1.record()
                torch.cuda.synchronize()
                if k >= warmup:
                    total += e0.elapsed_time(e1) / 1e3
            t = total / steps
            x = xbuf.view(batch, lb.batch_stride)[:, ib].reshape(batch, n, nrhs)
            r = b - torch.bmm(a, x)
            bwd = (torch.linalg.matrix_norm(r) /
                   (torch.linalg.matrix_norm(a) * torch.linalg.matrix_norm(x) + torch.linalg.matrix_norm(b)))
            entry = {"solves_per_s": batch / t, "ms_per_step": t * 1e3,
                     "max_solve_backward_error": float(bwd.max()), "info_nonzero": int((info != 0).sum())}
            if prec is not None:
                its = it.cpu().numpy()
                entry["refinement_steps"] = {str(v): int((its == v).sum()) for v in sorted(set(its.tolist()))}
                entry["fell_back_to_fp64"] = int((its < 0).sum())
                entry["by_kappa_decade"] = {
                    f"1e{d}": {"instances": int((kdec == d).sum()),
                               "fallback": int(((kdec == d) & (its < 0)).sum()),
                               "mean_steps_converged": (float(its[(kdec == d) & (its >= 0)].mean())
                                                        if ((kdec == d) & (its >= 0)).any() else None)}
                    for d in range(2, 9) if (kdec == d).any()}
                entry["speedup_vs_fp64"] = res["fp64"]["ms_per_step"] / (t * 1e3)
            else:
                flops = n ** 3 / 3 if spd else 2 * n ** 3 / 3
                flops += 2 * n * n * nrhs
                entry["fp64_tflops"] = flops * batch / t / 1e12
                entry["fp64_frac_of_peak"] = entry["fp64_tflops"] / PEAK_FP64_TFLOPS
            if prec is None and not spd:
                # pivots of the FP64 LU vs the oracle on the best-conditioned instances
                order = torch.argsort(kappa).cpu().numpy()[:8]
                lo = Layout.make(n, n, panel=False)
                host = a[order].transpose(1, 2).contiguous().cpu().numpy().reshape(len(order), -1)
                hbuf = np.zeros(len(order) * lo.batch_stride)
                hbuf.reshape(len(order), lo.batch_stride)[:, :n * n] = host
                piv_ref, _ = oracle.getrf(lo, hbuf, len(order))
                piv = ipiv.view(batch, n)[torch.from_numpy(order).cuda()].cpu().numpy()
                entry["pivots_equal_oracle"] = f"{int((piv == piv_ref).all(axis=1).sum())}/{len(order)} " \
                                               "best-conditioned instances"
            res[name] = entry
        out[variant] = res
        del a, b, a0, abuf, lbuf, bbuf, xbuf
    return out


NSTREAMS = 8


def adjoint_workload(steps, warmup, batch=4096, nmin=32, nmax=192, seed=20240611):
    """Adjoint of the batched SPD solve x = A^{-1} b: n ~ U{32..192} per
    instance, A = G G^T + n I, b and the cotangent xbar ~ N(0,1), one RHS.
    Instances are grouped by n (one uniform batched call per size); the
    per-size call sequence is timed eagerly and replayed from a captured
    hipGraph.  FP64: forward dposv (keeps L) + backward from L, and the
    recompute mode (backward from A, L recomputed into Abar's buffer); FP32:
    forward sposv (FP32 storage, FP64 arithmetic on the AUTO path) and
    backward from L.  A sample per size is checked against the FP64
    reference adjoint (Abar = -sym(bbar x^T) folded to the lower entries)."""
    import torch

    from blast_amd import capi
    from blast_amd.layout import Layout
    from oracle import adjoint

    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=batch)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    groups = []
    for n in np.unique(sizes):
        n, cnt = int(n), int((sizes == n).sum())
        g = torch.randn(cnt, n, n, dtype=torch.float64, device="cuda", generator=gen)
        a = torch.bmm(g, g.transpose(1, 2).contiguous()) + n * torch.eye(n, dtype=torch.float64, device="cuda")
        bv = torch.randn(cnt, n, 1, dtype=torch.float64, device="cuda", generator=gen)
        xb = torch.randn(cnt, n, 1, dtype=torch.float64, device="cuda", generator=gen)
        ent = {"n": n, "count": cnt, "a": a, "b": bv, "xb": xb}
        for dt, tdt in ((np.float64, torch.float64), (np.float32, torch.float32)):
            la, lb = Layout.make(n, n, dt), Layout.make(n, 1, dt)
            ia = torch.from_numpy(la.index_map().reshape(-1)).cuda()
            ib = torch.from_numpy(lb.index_map().reshape(-1)).cuda()

            def pack(x, lay, idx):
                buf = torch.zeros(cnt * lay.batch_stride, dtype=tdt, device="cuda")
                # index_map: storage offset of logical (i, j), in row-major (i, j) order
                buf.view(cnt, lay.batch_stride).index_copy_(1, idx, x.to(tdt).reshape(cnt, -1))
                return capi.Operand(lay, buf)

            ops = {"A": pack(a, la, ia), "B": pack(bv, lb, ib), "Xb": pack(xb, lb, ib)}
            ops["L"] = capi.Operand(la, torch.zeros_like(ops["A"].tensor))
            ops["X"] = capi.Operand(lb, torch.zeros_like(ops["B"].tensor))
            ops["Ab"] = capi.Operand(la, torch.zeros_like(ops["A"].tensor))
            ops["Bb"] = capi.Operand(lb, torch.zeros_like(ops["B"].tensor))
            ops["ia"], ops["ib"] = ia, ib
            ent[dt.__name__] = ops
        groups.append(ent)
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    # the forward solves (fused posv, FP64 and FP32) give L and x
    for ent in groups:
        for dt in ("float64", "float32"):
            o = ent[dt]
            capi.posv(o["A"], o["L"], o["B"], o["X"], ent["count"], info)
            assert int(info[:ent["count"]].abs().sum()) == 0
    torch.cuda.synchronize()

    # the size groups are spread over a few streams (default priority), forked
    # from and joined back into the current stream, so small groups overlap
    streams = [torch.cuda.Stream() for _ in range(NSTREAMS)]

    def run(kind):
        cur = torch.cuda.current_stream()
        for st in streams:
            st.wait_stream(cur)
        for g, ent in enumerate(groups):
            c, sp = ent["count"], streams[g % NSTREAMS].cuda_stream
            if kind in ("fp64_forward", "fp32_forward"):
                o = ent["float64" if kind == "fp64_forward" else "float32"]
                capi.posv(o["A"], o["L"], o["B"], o["X"], c, stream=sp)
            elif kind == "fp64_backward":
                o = ent["float64"]
                capi.posv_backward(0, o["L"], o["X"], o["Xb"], o["Ab"], o["Bb"], c, stream=sp)
            elif kind == "fp64_backward_recompute":
                o = ent["float64"]
                capi.posv_backward(1, o["A"], o["X"], o["Xb"], o["Ab"], o["Bb"], c, stream=sp)
            else:
                o = ent["float32"]
                capi.posv_backward(0, o["L"], o["X"], o["Xb"], o["Ab"], o["Bb"], c, stream=sp)
        for st in streams:
            cur.wait_stream(st)

    flops = float(sum(e["count"] * (e["n"] ** 3 / 3 + 2 * e["n"] ** 2) for e in groups))
    res = {"batch": batch, "n": f"uniform {nmin}..{nmax}", "distinct_sizes": len(groups), "nrhs": 1,
           "streams": NSTREAMS}
    for kind in ("fp64_forward", "fp32_forward", "fp64_backward", "fp64_backward_recompute", "fp32_backward"):
        t_eager = time_steps(lambda: run(kind), steps, warmup)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            run(kind)
            s.synchronize()
            with torch.cuda.graph(graph, stream=s):
                run(kind)
        torch.cuda.current_stream().wait_stream(s)
        t_graph = time_steps(graph.replay, steps, warmup)
        res[kind] = {"instances_per_s_graph": batch / t_graph, "instances_per_s_eager": batch / t_eager,
                     "ms_per_step_graph": t_graph * 1e3, "ms_per_step_eager": t_eager * 1e3}
        if kind == "fp64_forward":
            res[kind]["fp64_tflops_graph"] = flops / t_graph / 1e12
            # the same forward solves as ONE persistent launch over all sizes
            # (blast_hip_dposv_uneven_batched, size-ordered queue)
            inst = (capi.SolveInstance * batch)()
            nlist, k = [], 0
            for ent in groups:
                o = ent["float64"]
                la, lb = o["A"].layout, o["B"].layout
                for i in range(ent["count"]):
                    inst[k].a = la.descriptor(o["A"].tensor.data_ptr() + 8 * i * la.batch_stride)
                    inst[k].l = la.descriptor(o["L"].tensor.data_ptr() + 8 * i * la.batch_stride)
                    inst[k].b = lb.descriptor(o["B"].tensor.data_ptr() + 8 * i * lb.batch_stride)
                    inst[k].x = lb.descriptor(o["X"].tensor.data_ptr() + 8 * i * lb.batch_stride)
                    nlist.append(ent["n"])
                    k += 1
            dinst = torch.from_numpy(np.frombuffer(bytes(inst), dtype=np.uint8).copy()).cuda()
            order = np.zeros(batch, dtype=np.int32)
            capi.check("blast_hip_uneven_order", capi.lib().blast_hip_uneven_order(
                batch, np.array(nlist, dtype=np.int32).ctypes.data, order.ctypes.data))
            queue = torch.from_numpy(np.concatenate([[0], order]).astype(np.int32)).cuda()
            uinfo = torch.zeros(batch, dtype=torch.int32, device="cuda")

            def one_launch():
                capi.check("blast_hip_dposv_uneven_batched", capi.lib().blast_hip_dposv_uneven_batched(
                    torch.cuda.current_stream().cuda_stream, dinst.data_ptr(), batch, queue.data_ptr(),
                    uinfo.data_ptr()))

            t_one = time_steps(one_launch, steps, warmup)
            assert int(uinfo.abs().sum()) == 0
            res["fp64_forward_uneven_one_launch"] = {
                "instances_per_s": batch / t_one, "ms_per_step": t_one * 1e3, "fp64_tflops": flops / t_one / 1e12,
                "kernel": "posv_uneven_kernel (one persistent launch, queue ordered by block count)"}
        # check a sample of every 16th size against the FP64 reference adjoint
        if kind.endswith("backward") or kind.endswith("recompute"):
            worst = 0.0
            for ent in groups[::16]:
                o = ent["float32" if kind == "fp32_backward" else "float64"]
                x = o["X"].logical(ent["count"])[0].astype(np.float64)
                ab_ref, bb_ref = adjoint.posv_backward(np.tril(ent["a"][0].cpu().numpy()), x,
                                                       ent["xb"][0].cpu().numpy())
                worst = max(worst, np.abs(o["Ab"].logical(ent["count"])[0] - ab_ref).max() / np.abs(ab_ref).max(),
                            np.abs(o["Bb"].logical(ent["count"])[0] - bb_ref).max() / np.abs(bb_ref).max())
            res[kind]["max_rel_err_sample_vs_fp64_reference"] = float(worst)
        if kind == "fp64_backward":
            # the same adjoints as ONE persistent launch over all sizes
            # (blast_hip_dposv_backward_uneven_batched, from L and x of the
            # forward), into separate outputs, checked against the per-size ones
            ainst = (capi.AdjointInstance * batch)()
            nlist, k, outs = [], 0, []
            for ent in groups:
                o = ent["float64"]
                la, lb = o["A"].layout, o["B"].layout
                ab2, bb2 = torch.zeros_like(o["Ab"].tensor), torch.zeros_like(o["Bb"].tensor)
                outs.append((o, ab2, bb2))
                for i in range(ent["count"]):
                    ainst[k].l = la.descriptor(o["L"].tensor.data_ptr() + 8 * i * la.batch_stride)
                    ainst[k].x = lb.descriptor(o["X"].tensor.data_ptr() + 8 * i * lb.batch_stride)
                    ainst[k].xb = lb.descriptor(o["Xb"].tensor.data_ptr() + 8 * i * lb.batch_stride)
                    ainst[k].ab = la.descriptor(ab2.data_ptr() + 8 * i * la.batch_stride)
                    ainst[k].bb = lb.descriptor(bb2.data_ptr() + 8 * i * lb.batch_stride)
                    nlist.append(ent["n"])
                    k += 1
            dainst = torch.from_numpy(np.frombuffer(bytes(ainst), dtype=np.uint8).copy()).cuda()
            aorder = np.zeros(batch, dtype=np.int32)
            capi.check("blast_hip_uneven_order", capi.lib().blast_hip_uneven_order(
                batch, np.array(nlist, dtype=np.int32).ctypes.data, aorder.ctypes.data))
            aqueue = torch.from_numpy(np.concatenate([[0], aorder]).astype(np.int32)).cuda()
            ainfo = torch.zeros(batch, dtype=torch.int32, device="cuda")

            def one_backward():
                capi.check("blast_hip_dposv_backward_uneven_batched",
                           capi.lib().blast_hip_dposv_backward_uneven_batched(
                               torch.cuda.current_stream().cuda_stream, dainst.data_ptr(), batch, aqueue.data_ptr(),
                               ainfo.data_ptr()))

            t_one = time_steps(one_backward, steps, warmup)
            assert int(ainfo.abs().sum()) == 0
            diff = max(max(float((ab2 - o["Ab"].tensor).abs().max() / o["Ab"].tensor.abs().max()),
                           float((bb2 - o["Bb"].tensor).abs().max() / o["Bb"].tensor.abs().max()))
                       for o, ab2, bb2 in outs)
            res["fp64_backward_uneven_one_launch"] = {
                "instances_per_s": batch / t_one, "ms_per_step": t_one * 1e3,
                "max_rel_diff_vs_per_size_backward": diff,
                "kernel": "posv_backward_uneven_kernel (one persistent launch, queue ordered by block count)"}
    return res


def cpu_baseline(la, lb, abuf, bbuf, target_s):
    """BLAST-restated CPU Cholesky solve (oracle/blast_avx2.cpp: BLAST's panel
    potrf on AVX2 register tiles of 12/8/4 x 4, its scalar trsm; bit-identical
    to the scalar oracle) on the host cores, one instance per core, on the
    first 64 instances of the batch repeated for about target_s seconds (a
    2048-instance chunk, 256 MiB, reused per call)."""
    from oracle import oracle

    sample, copies = 64, 32
    a = abuf[: sample * la.batch_stride].cpu().numpy()
    b = bbuf[: sample * lb.batch_stride].cpu().numpy()
    ac, bc = np.tile(a, copies), np.tile(b, copies)
    lc, xc = np.zeros_like(ac), np.zeros_like(bc)
    cores = min(16, len(os.sched_getaffinity(0)))
    chunk = sample * copies
    oracle.posv_avx2_mt(la, ac, la, lc, lb, bc, lb, xc, chunk, threads=cores)  # warm-up and calibration
    t0 = time.perf_counter()
    oracle.posv_avx2_mt(la, ac, la, lc, lb, bc, lb, xc, chunk, threads=cores)
    calls = max(1, min(400, int(target_s / max(time.perf_counter() - t0, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(calls):
        _, used = oracle.posv_avx2_mt(la, ac, la, lc, lb, bc, lb, xc, chunk, threads=cores)
    dt = time.perf_counter() - t0
    rate = chunk * calls / dt
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": rate,
        "unit": "solves/s",
        "cores": used,
        "kind": "port",
        "label": f"BLAST-restated CPU ({used} cores)",
        "gflops_per_core": rate * FLOPS_PER_INSTANCE / used / 1e9,
        "cpu_model": cpu_model,
        "sample": f"{chunk * calls} instances ({calls} calls on the first {sample} instances of the batch x {copies}), "
                  f"BLAST's panel potrf on AVX2 register tiles + its scalar trsm, one instance per core, {dt:.1f} s",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); every instance is independent, "
              "multi-GPU sharding is a contiguous split of the batch", file=sys.stderr)
        sys.exit(2)
    import torch

    from blast_amd import capi

    if args.calibrate_copy:
        src = torch.ones(256 * 2 ** 20 // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.only in ("gemm", "riccati", "mixed256", "adjoint", "convert", "blocked", "blocked_lu"):
        res = {"gemm": lambda: gemm_workload(args.steps, args.warmup),
               "riccati": lambda: riccati_workload(args.steps, args.warmup, args.riccati_count),
               "mixed256": lambda: mixed256_workload(args.steps, args.warmup),
               "adjoint": lambda: adjoint_workload(args.steps, args.warmup),
               "convert": lambda: convert_workload(args.steps, args.warmup),
               "blocked": lambda: blocked_workload(args.steps, args.warmup),
               "blocked_lu": lambda: blocked_lu_workload(args.steps, args.warmup)}[args.only]()
        print(json.dumps({args.only: res}))
        return
    batch = args.batch
    la, lb, abuf, bbuf = make_inputs(batch)
    lbuf = torch.zeros_like(abuf)
    xbuf = torch.zeros_like(bbuf)
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    A, B = capi.Operand(la, abuf), capi.Operand(lb, bbuf)
    L, X = capi.Operand(la, lbuf), capi.Operand(lb, xbuf)
    capi.set_algorithm(capi.ALGO_AUTO)

    def plain():
        capi.posv(A, L, B, X, batch, info)

    t_plain = time_steps(plain, args.steps, args.warmup)
    se, fe = backward_errors(la, lb, abuf, xbuf, bbuf, batch, lbuf)
    assert int(info.abs().sum()) == 0
    plain_rate = batch / t_plain
    result = {
        "plain_fp64": {
            "solves_per_s": plain_rate,
            "ms_per_step": t_plain * 1e3,
            "tflops": FLOPS_PER_INSTANCE * batch / t_plain / 1e12,
            "max_solve_backward_error": float(se.max()),
            "max_potrf_backward_error": float(fe.max()),
            "contract": "dposv: A (lower) read, L and X written",
        }
    }
    # Mixed precision + FP64 refinement (dsposv contract: A read, X written;
    # A changes only for instances that fall back to FP64, none here).
    it = torch.zeros(batch, dtype=torch.int32, device="cuda")
    # tol: stop when ||b - A x||_inf <= tol ||A||_inf ||x||_inf (the 1e-12
    # backward-error bar with margin); tol = 0 is LAPACK dsposv's eps*sqrt(n).
    for name, prec, tol in (("mixed_fp32_ir", 0, 1e-13), ("mixed_bf16_ir", 1, 1e-13),
                            ("mixed_fp32_ir_lapack_test", 0, 0.0)):
        xm = torch.zeros_like(bbuf)
        Xm = capi.Operand(lb, xm)

        def mixed():
            capi.sposv(prec, A, B, Xm, batch, it, info, tol=tol)

        t_m = time_steps(mixed, args.steps, args.warmup)
        sem, _ = backward_errors(la, lb, abuf, xm, bbuf, batch)
        its = it.cpu().numpy()
        result[name] = {
            "solves_per_s": batch / t_m,
            "ms_per_step": t_m * 1e3,
            "speedup_vs_plain_fp64": t_plain / t_m,
            "max_solve_backward_error": float(sem.max()),
            "refinement_steps": {str(k): int((its == k).sum()) for k in sorted(set(its.tolist()))},
            "fell_back_to_fp64": int((its < 0).sum()),
            "stopping_tol": tol if tol > 0 else "eps*sqrt(n)",
            "contract": "dsposv: A read, X written",
        }
    # value: the faster of the two paths the metric compares, among those
    # whose every solve meets the 1e-12 backward-error bar
    mixed = result["mixed_fp32_ir"]
    plain_ok = result["plain_fp64"]["max_solve_backward_error"] <= 1e-12
    mixed_ok = mixed["max_solve_backward_error"] <= 1e-12 and mixed["fell_back_to_fp64"] == 0
    if mixed_ok and (not plain_ok or mixed["solves_per_s"] > plain_rate):
        value, t_step, dtype = mixed["solves_per_s"], mixed["ms_per_step"] / 1e3, "f32 factor + f64 refinement"
        headline = "mixed_fp32_ir"
    else:
        value, t_step, dtype = plain_rate, t_plain, "f64"
        headline = "plain_fp64"

    # roofline of the kernel behind `value`; the plain FP64 kernel's beside it
    t_mixed = mixed["ms_per_step"] / 1e3
    roof_plain = {
        "bound": "hbm", "kernel": "posv_ll_kernel<8>",
        "achieved": BYTES_PLAIN * batch / t_plain / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
        "frac": BYTES_PLAIN * batch / t_plain / 1e9 / PEAK_HBM_GBS,
        "traffic": profiled_traffic("posv_ll_kernel<8>"), "algorithmic_bytes_per_instance": BYTES_PLAIN,
        "fp64_tflops": FLOPS_PER_INSTANCE * batch / t_plain / 1e12,
        "fp64_frac_of_peak": FLOPS_PER_INSTANCE * batch / t_plain / 1e12 / PEAK_FP64_TFLOPS,
    }
    roof_mixed = {
        "bound": "hbm", "kernel": "sposv_ir_kernel<8, true, false, true>",
        "achieved": BYTES_SOLVE_ONLY * batch / t_mixed / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
        "frac": BYTES_SOLVE_ONLY * batch / t_mixed / 1e9 / PEAK_HBM_GBS,
        "traffic": profiled_traffic("sposv_ir_kernel<8, true, false, true>"),
        "algorithmic_bytes_per_instance": BYTES_SOLVE_ONLY,
        "timed": "one step = the mixed kernel + the FP64 fallback launch (exits at once when nothing fell back)",
    }
    if headline == "mixed_fp32_ir":
        roofline = dict(roof_mixed, plain_fp64=roof_plain)
    else:
        roofline = dict(roof_plain, mixed_fp32_ir=roof_mixed)

    line = {
        "metric": "batched potrf+trsm solves/s @n=128: mixed+IR vs plain FP64, both <=1e-12 bwd err",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic: A = G G^T + 128 I, G ~ N(0,1); B ~ U(-1,1); torch Philox seed 20240611",
        "config": {
            "workload": f"Cholesky solve potrf + trsm(L) + trsm(L^T), n={N}, nrhs={NRHS}, batch={batch}, "
                        "FP64, panel storage",
            "headline_path": headline,
            "paths": result,
        },
        "roofline": roofline,
    }
    if args.only == "posv":
        print(json.dumps(line))
        return
    line["config"]["other_workloads"] = {
        "dgemm_64_batch4096": gemm_workload(args.steps, args.warmup),
        "convert_n128_batch4096": convert_workload(args.steps, args.warmup),
        f"riccati_uneven_{args.riccati_count}_h10": riccati_workload(args.steps, args.warmup, args.riccati_count),
        "mixed_n256_nrhs8_batch2048": mixed256_workload(min(args.steps, 5), min(args.warmup, 1)),
        "adjoint_spd_solve_n32_192_batch4096": adjoint_workload(min(args.steps, 5), min(args.warmup, 1)),
        "blocked_n512_nrhs4_batch256": blocked_workload(min(args.steps, 5), min(args.warmup, 1)),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(la, lb, abuf, bbuf, args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
