#!/usr/bin/env python3
"""Soft Smith-Waterman (affine gap) forward+backward throughput on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload ragged|fixed|long|embed|nw|sv]

Workload "ragged" (default): 2048 pairs with L1, L2 drawn independently and
uniformly from [64, 1024], then each pair's N(0,1) FP32 scores in pair order
(torch.Generator seed 2); gap_open=-8, gap_extend=-0.5, temperature=0.5.
Workload "fixed": [256, 512, 512] scores N(0,1)*2 - 0.5 (seed 1), gap_open=-11,
gap_extend=-1, temperature=1.
Workload "long": [16, 4096, 4096] N(0,1) (seed 3), gap_open=-11, gap_extend=-1,
temperature=1; also the HVP along an N(0,1) direction and the step's peak HBM
against a reverse sweep that reads a full M/I/D table.
Workload "embed": 512 pairs of BF16 embeddings Q [M, 128], K [N, 128] with M,
N uniform in [128, 768] (seed 4); scores Q K^T by a BF16 GEMM with FP32
accumulation, the ragged sweep, dQ and dK through the map.
Workload "nw" / "sv": the ragged batch of "ragged" through the
Needleman-Wunsch (global) / Saigo-Vert affine sweeps: value, map and
parameter gradients.
Workload "grid": the reference's own Smith-Waterman benchmark grid (22 shapes
from 64 x 64 to 2048 x 2048, batches 1 to 128) through the public calls, with
that harness's inputs, parameters and timing protocol.

A step is one forward (value, log-partition) plus backward (marginal map and
the gradients of gap_open, gap_extend and temperature) over the whole batch:
one call of ohm_sw_affine_forward with every output requested, i.e. the
planning kernel, the forward sweep and the reverse sweep (or one launch doing
both per pair).  The launch shape is autotuned on the batch first
(orihime_amd.tune; --no-tune keeps the automatic choice); the results do not
depend on it.  The ragged batch runs as one packed persistent launch; it is
compared with the padded launch ([B, 1024, 1024] with the cells beyond each
pair's lengths masked, i.e. set to the reference's mask sentinel), which is
how the batch runs without ragged support.  GCUPS counts the useful cells
sum_b L1_b * L2_b for both.

Timed steps replay a HIP graph of one step on inputs already in HBM.  Kernel
durations come from HIP events recorded between the launches of the same
calls (ohm_set_stage_events) in a separate eager pass of the same K steps.

Roofline per cell, affine (algorithmic, max-shifted FP32 log-sum-exp):
forward 11 transcendentals and 4 bytes (read the score), reverse 11 and 4
bytes (write the map), a fused forward+reverse launch 22 and 8.  `roofline`
reports the dominant kernel against HBM (the contract's form) and, under
`binding`, against the transcendental issue rate measured on the device
(orihime_amd/csrc/peak_probe.hip); `traffic` is the dominant kernel's HBM
bytes per launch from the committed byte-counter profile of the workload.

The CPU baseline is the reference's CPU kernels (compiled from source into
oracle/_ref) or, where that build is absent, the oracle's C restatement of
them, timed on a bounded prefix of the same batch.

Prints one JSON line on stdout; progress goes to stderr.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "GCUPS fwd+bwd soft-SW-affine, ragged vs padded, % of roofline, 1 GPU"

# MI355X published / derived peaks
HBM_PEAK_GBS = 8000.0
CUS, CLOCK_HZ = 256, 2.4e9
VALU_LANE_OPS = CUS * 64 * CLOCK_HZ          # non-packed FP32 lane-ops/s (39.3 T)
TRANS_PEAK = VALU_LANE_OPS / 4               # v_exp/v_log/v_rcp at quarter rate (9.83 T/s)
TRANS_PER_CELL = {"forward": 11, "reverse": 11, "fused": 22}
BYTES_PER_CELL = {"forward": 4, "reverse": 4, "fused": 8}
KERNEL_NAMES = {
    "plan": "sw_plan_kernel",
    "forward": "sw_affine_scaled_forward_kernel<true, W, CK>",
    "reverse": "sw_affine_scaled_backward_kernel<W, CK>",
    # one launch running each pair's forward and reverse sweep (the launch
    # shape the autotuner picks on the ragged batch); W waves per workgroup,
    # CK steps between checkpoints
    "fused": "sw_affine_scaled_fused_kernel<W, CK>",
}

# The reference's Smith-Waterman shape catalog ("rectangular-production-v1",
# benchmarks/grid_cases.json:884-1306 of the reference): case id, batch, L1,
# L2 and the explicit lengths of the one ragged case (else every pair is
# full).  Kept as data so the benchmark needs no reference files.
RECT_CASES = [
    ("rect-b1-64x64", 1, 64, 64, None),
    ("rect-b1-128x128", 1, 128, 128, None),
    ("rect-b1-256x256", 1, 256, 256, None),
    ("rect-b1-512x512", 1, 512, 512, None),
    ("rect-b1-1024x1024", 1, 1024, 1024, None),
    ("rect-b1-2048x2048", 1, 2048, 2048, None),
    ("rect-b1-128x2048", 1, 128, 2048, None),
    ("rect-b1-2048x128", 1, 2048, 128, None),
    ("rect-b1-512x2048", 1, 512, 2048, None),
    ("rect-b1-2048x512", 1, 2048, 512, None),
    ("rect-b1-511x1023", 1, 511, 1023, None),
    ("rect-b1-1023x511", 1, 1023, 511, None),
    ("rect-b1-1023x2047", 1, 1023, 2047, None),
    ("rect-b1-2047x1023", 1, 2047, 1023, None),
    ("rect-b8-64x64", 8, 64, 64, None),
    ("rect-b32-64x64", 32, 64, 64, None),
    ("rect-b128-64x64", 128, 64, 64, None),
    ("rect-b8-128x128", 8, 128, 128, None),
    ("rect-b32-128x128", 32, 128, 128, None),
    ("rect-b8-256x256", 8, 256, 256, None),
    ("rect-b4-512x512", 4, 512, 512, None),
    ("rect-b4-2048x2048-ragged", 4, 2048, 2048,
     [[2048, 2048], [1537, 997], [511, 2047], [127, 257]]),
]
# Scoring parameters of that harness (benchmarks/grid.py:42-48).
GRID_PARAMS = {"sw": {"gap_score": -0.75, "temperature": 0.73},
               "sw_affine": {"gap_open_score": -1.25, "gap_extend_score": -0.25,
                             "temperature": 0.73}}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# --------------------------------------------------------------------------
# Workloads (host-generated, seeded)
# --------------------------------------------------------------------------


def ragged_pairs(torch, seed=2, B=2048, lo=64, hi=1024):
    g = torch.Generator().manual_seed(seed)
    lens = torch.randint(lo, hi + 1, (B, 2), generator=g)
    return [torch.randn(int(a), int(b), generator=g) for a, b in lens.tolist()]


def fixed_batch(torch, seed=1, B=256, L=512):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(B, L, L, generator=g) * 2 - 0.5


def long_batch(torch, seed=3, B=16, L=4096):
    """(scores, direction), both N(0,1), drawn in that order."""
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(B, L, L, generator=g)
    return x, torch.randn(B, L, L, generator=g)


def embed_inputs(torch, B=512, seed=4, lo=128, hi=768, D=128):
    """Q [B, hi, D], K [B, hi, D] BF16 N(0,1)/sqrt(D) (rows past each pair's
    length are drawn too and ignored), lengths [B, 2] int32 ~ U[lo, hi];
    lengths drawn first."""
    g = torch.Generator().manual_seed(seed)
    ln = torch.randint(lo, hi + 1, (B, 2), generator=g).to(torch.int32)
    Q = (torch.randn(B, hi, D, generator=g) / D ** 0.5).to(torch.bfloat16)
    K = (torch.randn(B, hi, D, generator=g) / D ** 0.5).to(torch.bfloat16)
    return Q, K, ln


# --------------------------------------------------------------------------
# HIP events through the runtime torch already loaded
# --------------------------------------------------------------------------


class Hip:
    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        self.lib.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        self.lib.hipEventRecord.argtypes = [vp, vp]
        self.lib.hipEventSynchronize.argtypes = [vp]
        self.lib.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        self.lib.hipEventDestroy.argtypes = [vp]

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def event(self) -> int:
        ev = ctypes.c_void_p()
        self._ok(self.lib.hipEventCreate(ctypes.byref(ev)), "hipEventCreate")
        return ev.value

    def record(self, ev: int, stream: int) -> None:
        self._ok(self.lib.hipEventRecord(ev, stream or None), "hipEventRecord")

    def elapsed_ms(self, a: int, b: int) -> float:
        self._ok(self.lib.hipEventSynchronize(b), "hipEventSynchronize")
        ms = ctypes.c_float()
        self._ok(self.lib.hipEventElapsedTime(ctypes.byref(ms), a, b), "hipEventElapsedTime")
        return float(ms.value)

    def destroy(self, ev: int) -> None:
        self.lib.hipEventDestroy(ev)


def measured_rates(torch, hip) -> dict:
    """Issue rates of the instruction kinds the sweeps are made of, in
    T lane-operations/s, from the probe kernels in libohm_probe.so."""
    lib = ctypes.CDLL(os.path.join(ROOT, "orihime_amd", "libohm_probe.so"))
    lib.ohm_probe_launch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p]
    lib.ohm_probe_ops_per_thread.argtypes = [ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int64)]
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    blocks, iters = cus * 8, 4096  # 8 waves per SIMD
    sink = torch.empty(blocks * 256, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    out = {}
    for name, kind in (("exp", 0), ("log", 1), ("fma", 2), ("dpp", 3), ("exp+4fma", 4)):
        ops = ctypes.c_int64()
        if lib.ohm_probe_ops_per_thread(kind, iters, ctypes.byref(ops)) != 0:
            raise RuntimeError("probe op count")
        best = None
        for _ in range(3):
            a, b = hip.event(), hip.event()
            hip.record(a, stream)
            if lib.ohm_probe_launch(kind, blocks, iters, sink.data_ptr(), stream) != 0:
                raise RuntimeError("probe launch failed")
            hip.record(b, stream)
            ms = hip.elapsed_ms(a, b)
            hip.destroy(a)
            hip.destroy(b)
            best = ms if best is None else min(best, ms)
        out[name] = round(ops.value * blocks * 256 / (best * 1e-3) / 1e12, 3)
    return out


# --------------------------------------------------------------------------
# One configured launch (packed or padded), with its outputs and workspace
# --------------------------------------------------------------------------


class Launch:
    def __init__(self, torch, lib, scores, lengths, offsets, L1max, L2max, params):
        self.torch, self.lib = torch, lib
        self.scores, self.lengths, self.offsets = scores, lengths, offsets
        self.L1max, self.L2max = L1max, L2max
        self.params = params
        B = lengths.shape[0] if lengths is not None else scores.shape[0]
        dev = scores.device
        self.B = B
        self.value = torch.empty(B, device=dev)
        self.map = torch.empty_like(scores)
        self.grads = [torch.empty(B, device=dev) for _ in range(3)]
        self.ws = None
        self.size_workspace()
        total = scores.numel() if offsets is not None else -1
        self.ws_value = torch.empty(lib.workspace_bytes(True, lib.OP_VALUE, B, L1max, L2max, total),
                                    dtype=torch.uint8, device=dev)
        self.value_only = torch.empty(B, device=dev)
        self.graph = None

    def size_workspace(self):
        """(Re)allocate the forward + reverse workspace for the checkpoint
        stride in effect (the tuner may change it)."""
        total = self.scores.numel() if self.offsets is not None else -1
        nbytes = self.lib.workspace_bytes(True, self.lib.OP_FORWARD, self.B, self.L1max,
                                          self.L2max, total)
        if self.ws is None or self.ws.numel() != nbytes:
            self.ws = None
            self.ws = self.torch.empty(nbytes, dtype=self.torch.uint8, device=self.scores.device)

    def step(self):
        go, ge, T = self.params
        self.lib.affine_forward(self.scores, self.lengths, self.offsets, go, ge, T,
                                value=self.value, map=self.map, grad_open=self.grads[0],
                                grad_ext=self.grads[1], grad_temp=self.grads[2],
                                L1max=self.L1max, L2max=self.L2max, workspace=self.ws)

    def forward_only(self):
        go, ge, T = self.params
        self.lib.affine_forward(self.scores, self.lengths, self.offsets, go, ge, T,
                                value=self.value_only, L1max=self.L1max, L2max=self.L2max,
                                workspace=self.ws_value)

    def outputs(self):
        return [self.value, self.map] + self.grads

    def capture(self):
        torch = self.torch
        self.graph = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.step()  # one eager call on the capture stream first
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        with torch.cuda.graph(self.graph):
            self.step()
        torch.cuda.synchronize()


def _profiled(name: str, tag: str, kernel: str, shape=None):
    """Per-launch figure of `kernel` from the `tag:` line of a committed
    counter profile: the entry of the launch shape (waves, checkpoint stride)
    when given and present, else the first entry of that kernel."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith(tag + ":"):
            table = json.loads(line.split(":", 1)[1])
            if shape is not None:
                exact = KERNEL_NAMES[kernel].replace("CK", str(shape[1])).replace("W", str(shape[0]))
                if exact in table:
                    return int(table[exact]), os.path.relpath(path, ROOT)
            key = KERNEL_NAMES[kernel].split("<")[0] + "<" + ("true" if kernel == "forward" else "")
            for k, v in table.items():
                if k.startswith(key):
                    return int(v), os.path.relpath(path, ROOT)
    return None, None


def profiled_traffic(workload: str, kernel: str, shape=None):
    """HBM bytes per launch of `kernel` from the committed byte-counter profile
    of this workload (rocprofv3 FETCH_SIZE / WRITE_SIZE passes, calibrated with
    a device copy of known size), or None when there is none."""
    return _profiled(f"sw_affine_{workload}_hbm_traffic.md", "traffic_bytes_per_launch", kernel, shape)


def profiled_valu(workload: str, kernel: str, shape=None):
    """Wave-level VALU instructions per launch of `kernel` from the committed
    SQ-counter profile of this workload (rocprofv3 SQ_INSTS_VALU), or None."""
    return _profiled(f"sw_affine_{workload}_instructions.md", "valu_instructions_per_launch", kernel, shape)


def timed(torch, fn, steps):
    """Mean ms per call over `steps` calls, bracketed by device syncs."""
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(steps):
        fn()
    end.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return start.elapsed_time(end) / steps, wall


def kernel_durations(torch, lib, hip, launch, steps):
    """Mean duration (ms) of each of the step's launches, from HIP events
    recorded on the launch stream between them."""
    stream = torch.cuda.current_stream().cuda_stream
    names = ("plan", "forward", "forward_fallback", "reverse", "reverse_fallback")
    evs = [[hip.event() for _ in range(len(names) + 1)] for _ in range(steps)]
    torch.cuda.synchronize()
    try:
        for e in evs:
            hip.record(e[0], stream)
            lib.set_stage_events(e[1:])
            launch.step()
    finally:
        lib.set_stage_events([])
    torch.cuda.synchronize()
    out = {k: 0.0 for k in names}
    for e in evs:
        for i, k in enumerate(names):
            out[k] += hip.elapsed_ms(e[i], e[i + 1])
        for x in e:
            hip.destroy(x)
    return {k: v / steps for k, v in out.items()}


# --------------------------------------------------------------------------
# CPU baseline: the reference CPU kernels on a bounded prefix of the batch
# --------------------------------------------------------------------------


def cpu_model() -> str:
    """The host CPU's model name (lscpu), or the platform's processor string."""
    import platform
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return platform.processor() or "unknown"


def host_threads() -> int:
    """Threads the CPU baseline may use: the host's cores, at most 16 (one
    GPU's share of a shared host)."""
    return max(1, min(16, os.cpu_count() or 1))


def _cpu_run(torch, pairs, params, seconds, family, threads):
    import numpy as np

    from oracle import oracle as O

    run = {"sw": O.sw_affine_forward, "nw": O.nw_affine_forward, "sv": O.sv_affine_forward}[family]
    built = {"sw": O.ref_available, "nw": O.nw_ref_available, "sv": O.sv_ref_available}[family]
    kind = "reference" if built() else "port"
    old = torch.get_num_threads()
    torch.set_num_threads(threads)
    go, ge, T = params
    cells = done = 0
    elapsed = 0.0
    chunk = max(1, threads)
    try:
        while done < len(pairs) and elapsed < seconds:
            part = pairs[done:done + chunk]
            L1 = max(p.shape[0] for p in part)
            L2 = max(p.shape[1] for p in part)
            pad = np.zeros((len(part), L1, L2), np.float32)
            ln = np.zeros((len(part), 2), np.int32)
            for i, p in enumerate(part):
                pad[i, :p.shape[0], :p.shape[1]] = p
                ln[i] = p.shape
            t0 = time.perf_counter()
            if kind == "reference":
                run(pad, go, ge, T, ln, impl="ref")
            else:  # the single-threaded restatement: threads over pairs
                O.over_pairs(run, list(part), go, ge, T, threads=threads, chunk=1)
            elapsed += time.perf_counter() - t0
            cells += int((ln[:, 0].astype(np.int64) * ln[:, 1]).sum())
            done += len(part)
    finally:
        torch.set_num_threads(old)
    return kind, cells, done, elapsed


def cpu_baseline(torch, pairs, params, seconds, family="sw"):
    """The reference's CPU kernels (oracle/_ref; the restatement when they are
    not built) on the first pairs of the batch: about `seconds` of work on the
    host's threads (at most 16), then a quarter of that on one thread.
    family: "sw" (Smith-Waterman affine), "nw" or "sv" (affine)."""
    threads = host_threads()
    kind, cells, done, elapsed = _cpu_run(torch, pairs, params, seconds, family, threads)
    _, cells1, done1, elapsed1 = _cpu_run(torch, pairs, params, seconds / 4, family, 1)
    what = "forward+backward (value, map, gap and temperature gradients)"
    return {
        "value": round(cells / elapsed / 1e9, 6),
        "unit": "GCUPS",
        "cores": threads,
        "threads": threads,
        "kind": kind,
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
        "sample": f"first {done} of {len(pairs)} pairs ({cells} cells), {what}, {elapsed:.1f} s",
        "single_thread": {
            "value": round(cells1 / elapsed1 / 1e9, 6), "unit": "GCUPS", "threads": 1,
            "sample": f"first {done1} pairs ({cells1} cells), {what}, {elapsed1:.1f} s"},
    }


def hbm_copy_peak(torch, hip, nbytes=2 << 30, reps=10) -> float:
    """HBM bandwidth of a device-to-device copy on this GPU (GB/s, bytes read
    plus bytes written), best of `reps`."""
    src = torch.ones(nbytes // 4, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream().cuda_stream
    dst.copy_(src)
    torch.cuda.synchronize()
    best = None
    for _ in range(reps):
        a, b = hip.event(), hip.event()
        hip.record(a, stream)
        dst.copy_(src)
        hip.record(b, stream)
        ms = hip.elapsed_ms(a, b)
        hip.destroy(a)
        hip.destroy(b)
        best = ms if best is None else min(best, ms)
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * nbytes / (best * 1e-3) / 1e9


def oracle_errors(pairs, outputs, params, n_values=256, n_maps=32):
    """Errors of a packed launch's results against the FP64 oracle on pairs
    spread over the size range (outside any timed region).  outputs: value
    [B] and the per-pair maps (numpy)."""
    import numpy as np

    from oracle import oracle as O

    value, maps = outputs
    order = np.argsort([p.size for p in pairs], kind="stable")
    pick = lambda n: sorted(int(order[k]) for k in
                            np.linspace(0, len(order) - 1, min(n, len(order))).round().astype(int))
    iv, im = pick(n_values), pick(n_maps)
    go, ge, T = params
    f64v = np.asarray(O.over_pairs(O.sw_affine_value, [pairs[i] for i in iv], go, ge, T,
                                   prec="f64", threads=host_threads(), chunk=4)[0])
    got = np.asarray(value, np.float64)[iv]
    vrel = float(np.max(np.abs(got - f64v) / np.maximum(np.abs(f64v), 1e-3)))
    _, f64m, *_ = O.over_pairs(O.sw_affine_forward, [pairs[i] for i in im], go, ge, T,
                               prec="f64", threads=host_threads(), chunk=1)
    mabs = max(float(np.max(np.abs(maps[i] - m), initial=0.0)) for i, m in zip(im, f64m))
    return {"value_rel_max": vrel, "map_abs_max": mabs, "values_checked": len(iv),
            "maps_checked": len(im), "against": "FP64 oracle",
            "pairs": "spread over the size range (evenly spaced ranks of cell count)"}


def _padded_oracle_errors(batch, outputs, params, idx):
    """Value / map errors of a padded launch against the FP64 oracle."""
    import numpy as np

    from oracle import oracle as O

    go, ge, T = params
    sub = batch[idx].astype(np.float64)
    f64 = O.over_pairs(O.sw_affine_forward, [x for x in sub], go, ge, T, prec="f64",
                       threads=host_threads(), chunk=1)
    v, m = outputs
    vrel = float(np.max(np.abs(v[idx] - np.asarray(f64[0])) / np.maximum(np.abs(np.asarray(f64[0])), 1e-3)))
    mabs = float(max(np.max(np.abs(m[k] - f64[1][k])) for k in range(len(idx))))
    return {"value_rel_max": vrel, "map_abs_max": mabs, "values_checked": len(idx),
            "maps_checked": len(idx), "against": "FP64 oracle", "pairs": list(map(int, idx))}


def fixed_record(torch, lib, hip, args, dev):
    """[256, 512, 512] (N(0,1)*2 - 0.5, seed 1; go=-11, ge=-1, T=1): fwd+bwd
    step (graph replay), its kernel durations, sampled FP64 errors."""
    params = (-11.0, -1.0, 1.0)
    x = fixed_batch(torch)
    launch = Launch(torch, lib, x.to(dev), None, None, 512, 512, params)
    tuned = _tune(torch, launch, params, args)
    for _ in range(max(1, args.warmup)):
        launch.step()
    torch.cuda.synchronize()
    launch.capture()
    ms, _ = timed(torch, launch.graph.replay, args.steps)
    kd = kernel_durations(torch, lib, hip, launch, max(3, args.steps // 4))
    cells = x.numel()
    idx = [0, 97, 255]
    err = _padded_oracle_errors(x.numpy(), (launch.value.cpu().numpy(),
                                            launch.map[idx].cpu().numpy()), params, idx)
    log(f"fixed 256x512^2 fwd+bwd: {ms:.3f} ms/step, {cells / ms / 1e6:.1f} GCUPS")
    return {"workload": "fixed-shape affine soft-SW fwd+bwd, [256, 512, 512] N(0,1)*2-0.5 (seed 1), "
                        "go=-11 ge=-1 T=1",
            "cells": cells, "ms_per_step": round(ms, 4), "gcups": round(cells / ms / 1e6, 3),
            "kernel_ms": {k: round(v, 4) for k, v in kd.items()}, "hip_graph": True,
            "launch_config": tuned, "oracle_err": err}


def long_record(torch, lib, hip, args, dev):
    """[16, 4096, 4096] (N(0,1), seed 3; go=-11, ge=-1, T=1): fwd+bwd step
    with the checkpointed reverse, the HVP along an N(0,1) direction, peak
    HBM against a reverse that reads a full M/I/D table, FP64 errors of pair
    0 (value, central 64 x 64 map crop)."""
    import numpy as np

    from oracle import oracle as O

    params = (-11.0, -1.0, 1.0)
    x, V = long_batch(torch)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    launch = Launch(torch, lib, x.to(dev), None, None, 4096, 4096, params)
    tuned = _tune(torch, launch, params, args)
    for _ in range(max(1, args.warmup)):
        launch.step()
    torch.cuda.synchronize()
    steps = max(3, args.steps // 4)
    ms, _ = timed(torch, launch.step, steps)
    kd = kernel_durations(torch, lib, hip, launch, steps)
    peak = torch.cuda.max_memory_allocated() - base
    strips = 4096 // 64
    full_table = 16 * strips * (4096 + 63) * 64 * 3 * 4  # M, I, D per cell slot
    bnd = 16 * strips * 4096 * 8 * 4
    ckpt_ws = launch.ws.numel()
    cells = x.numel()
    Vd = V.to(dev)
    mt = torch.empty_like(launch.scores)
    tws = torch.empty(lib.workspace_bytes(True, lib.OP_TANGENT, 16, 4096, 4096, -1),
                      dtype=torch.uint8, device=dev)
    hvp = lambda: lib.affine_tangent(launch.scores, None, None, *params, direction=Vd,
                                     map_tangent=mt, workspace=tws)
    hvp()
    torch.cuda.synchronize()
    hms, _ = timed(torch, hvp, steps)
    sl = slice(2016, 2080)
    x0 = x[0].numpy().astype(np.float64)
    f64 = O.sw_affine_forward(x0[None], *params, prec="f64")
    v0 = float(launch.value[0])
    crop = launch.map[0, sl, sl].cpu().numpy()
    err = {"value_rel_max": abs(v0 - float(f64[0][0])) / abs(float(f64[0][0])),
           "map_abs_max": float(np.max(np.abs(crop - f64[1][0, sl, sl]))),
           "values_checked": 1, "maps_checked": "pair 0, central 64 x 64 crop",
           "against": "FP64 oracle"}
    del Vd, mt, tws, launch
    torch.cuda.empty_cache()
    log(f"long 16x4096^2 fwd+bwd: {ms:.3f} ms/step, HVP {hms:.3f} ms, peak HBM {peak / 1e9:.2f} GB")
    return {"workload": "long pairs affine soft-SW, [16, 4096, 4096] N(0,1) (seed 3), go=-11 "
                        "ge=-1 T=1: fwd+bwd with the checkpointed reverse, and the HVP",
            "cells": cells, "ms_per_step": round(ms, 4), "gcups": round(cells / ms / 1e6, 3),
            "kernel_ms": {k: round(v, 4) for k, v in kd.items()}, "launch_config": tuned,
            "hvp_ms": round(hms, 4), "hvp_gcups": round(cells / hms / 1e6, 3),
            "peak_hbm_bytes": int(peak), "workspace_bytes": int(ckpt_ws),
            "full_table_workspace_bytes": int(full_table + bnd),
            "full_table_peak_bytes": int(peak - ckpt_ws + full_table + bnd),
            "oracle_err": err}


def _embed_parts(torch, hip, Q, K, lnd, params, steps):
    """Mean ms of the embedding-scored step's parts, each timed alone with
    HIP events: the score kernel (ragged MFMA) and the sweeps on its output
    (value + map); for comparison the sweeps that generate the scores
    themselves (``fused=True``); dQ / dK (ragged MFMA, transposed copies
    included)."""
    from orihime_amd import _lib, embed

    stream = torch.cuda.current_stream().cuda_stream
    go, ge, T = params
    M, N = Q.shape[1], K.shape[1]
    out = {}
    with torch.no_grad():
        scores, offsets = embed.pair_scores(Q, K, lnd)
        _, post = embed._sweep(scores, lnd, offsets, M, N, go, ge, T)
        gv = torch.ones(Q.shape[0], device=Q.device)
        dQ, dK = torch.empty_like(Q), torch.empty_like(K)
        parts = {
            "fused_sweeps": lambda: embed._fused_sweep(Q, K, lnd, go, ge, T),
            "scores": lambda: _lib.embed_scores(Q, K, lnd, offsets, scores),
            "sweeps": lambda: embed._sweep(scores, lnd, offsets, M, N, go, ge, T),
            "dq_dk": lambda: _lib.embed_grads(Q, K, lnd, offsets, post, gv, dQ, dK),
        }
        for name, fn in parts.items():
            fn()
            torch.cuda.synchronize()
            a, b = hip.event(), hip.event()
            hip.record(a, stream)
            for _ in range(steps):
                fn()
            hip.record(b, stream)
            out[name] = round(hip.elapsed_ms(a, b) / steps, 4)
            hip.destroy(a)
            hip.destroy(b)
    return out


def embed_record(torch, hip, args, dev):
    """Embedding-scored alignment end to end: Q [512, M, 128], K [512, N, 128]
    BF16 N(0,1)/sqrt(128), M, N ~ U[128, 768] (seed 4); scores Q K^T (BF16
    products, FP32 accumulation), ragged soft-SW affine (go=-4, ge=-0.5,
    T=0.25), dQ and dK through the map.  A step is the forward and the
    backward."""
    import numpy as np

    from oracle import oracle as O
    from orihime_amd.embed import soft_align_embeddings

    params = (-4.0, -0.5, 0.25)
    Qc, Kc, ln = embed_inputs(torch)
    Q, K, lnd = Qc.to(dev), Kc.to(dev), ln.to(dev)
    cells = int((ln[:, 0].long() * ln[:, 1].long()).sum())
    Q.requires_grad_(True)
    K.requires_grad_(True)

    def step(fused=False):
        Q.grad = None
        K.grad = None
        soft_align_embeddings(Q, K, lnd, *params, fused=fused).sum().backward()

    from orihime_amd import embed, tune
    tuned = "automatic"
    tune.reset()
    if not args.no_tune:
        with torch.no_grad():
            scores, offsets = embed.pair_scores(Q, K, lnd)
            cfg = tune.autotune(scores, *params, lengths=lnd, offsets=offsets)
            del scores, offsets
        tuned = {"waves": cfg.waves, "groups_per_cu": cfg.groups_per_cu,
                 "teams": {0: "auto", 1: "on", 2: "off", 3: "grid", 4: "fused"}[cfg.teams],
                 "checkpoint_stride": cfg.ckpt or 8}
    for _ in range(max(1, args.warmup)):
        step()
        step(True)
    torch.cuda.synchronize()
    ms, _ = timed(torch, step, args.steps)
    ms_fused, _ = timed(torch, lambda: step(True), args.steps)
    parts = _embed_parts(torch, hip, Q.detach(), K.detach(), lnd, params, max(3, args.steps // 2))
    # FP64 errors of 16 pairs spread over the size range
    Q.grad = None
    K.grad = None
    v = soft_align_embeddings(Q, K, lnd, *params)
    v.sum().backward()
    vh, dQ, dK = v.detach().cpu().numpy(), Q.grad.float().cpu().numpy(), K.grad.float().cpu().numpy()
    lnh = ln.numpy()
    order = np.argsort(lnh[:, 0].astype(np.int64) * lnh[:, 1], kind="stable")
    idx = sorted(int(order[k]) for k in np.linspace(0, len(order) - 1, 16).round().astype(int))
    Q64, K64 = Qc.double().numpy(), Kc.double().numpy()
    S = [Q64[b, :lnh[b, 0]] @ K64[b, :lnh[b, 1]].T for b in idx]
    val, mp, *_ = O.over_pairs(O.sw_affine_forward, S, *params, prec="f64",
                               threads=host_threads(), chunk=1)
    vrel = float(np.max(np.abs(vh[idx] - np.asarray(val)) / np.abs(np.asarray(val))))
    gq = gk = scale = 0.0
    for k, b in enumerate(idx):
        m, n = lnh[b]
        tq, tk = mp[k] @ K64[b, :n], mp[k].T @ Q64[b, :m]
        scale = max(scale, float(np.abs(tq).max()), float(np.abs(tk).max()))
        gq = max(gq, float(np.abs(dQ[b, :m] - tq).max()))
        gk = max(gk, float(np.abs(dK[b, :n] - tk).max()))
    gcups = cells / ms / 1e6
    log(f"embedding-scored fwd+bwd: {ms:.3f} ms/step ({ms_fused:.3f} with the scores generated "
        f"inside the sweeps), {gcups:.1f} GCUPS; parts {parts}")
    return {"workload": "embedding-scored ragged soft-SW affine, Q/K [512, <=768, 128] BF16, "
                        "M,N ~ U[128,768] (seed 4), go=-4 ge=-0.5 T=0.25: ragged MFMA scores, "
                        "packed sweeps, dQ/dK",
            "cells": cells, "ms_per_step": round(ms, 4), "gcups": round(gcups, 3),
            "ms_per_step_scores_inside_sweeps": round(ms_fused, 4),
            "kernel_ms": parts, "launch_config": tuned,
            "gemm_gflop_per_step_useful": round(3 * 2.0 * 128 * cells / 1e9, 2),
            "oracle_err": {"value_rel_max": vrel, "dq_abs_max": gq, "dk_abs_max": gk,
                           "grad_scale": scale, "pairs_checked": len(idx),
                           "against": "FP64 oracle on the same BF16 values"}}


def embed_workload(torch, args, dev):
    """``--workload embed``: the embedding-scored record as its own line."""
    rec = embed_record(torch, Hip(), args, dev)
    line = {
        "metric": "GCUPS embedding-scored soft-SW-affine fwd+bwd (Q K^T, sweeps, dQ/dK), 1 GPU",
        "value": rec["gcups"], "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": rec["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 GEMM inputs, f32 sweeps",
        "data": "synthetic", "config": rec, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def _tune(torch, launch, params, args):
    if args.no_tune:
        return "automatic"
    from orihime_amd import tune
    t0 = time.perf_counter()
    cfg = tune.autotune(launch.scores, *params, lengths=launch.lengths, offsets=launch.offsets)
    launch.size_workspace()
    from orihime_amd import _lib
    return {"waves": cfg.waves, "groups_per_cu": cfg.groups_per_cu,
            "teams": {0: "auto", 1: "on", 2: "off", 3: "grid", 4: "fused"}[cfg.teams],
            "checkpoint_stride": _lib.checkpoint_stride(),
            "tuning_s": round(time.perf_counter() - t0, 2)}


def nw_traffic(kernel: str, family: str = "nw"):
    """HBM bytes per launch of the NW (SV) forward or reverse sweep from the
    committed FETCH_SIZE / WRITE_SIZE profile of the workload, or None."""
    path = os.path.join(ROOT, "profiles", f"{family}_affine_ragged2048_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            table = json.loads(line.split(":", 1)[1])
            key = (f"{family}_affine_forward_kernel<true" if kernel == "forward"
                   else f"{family}_affine_backward_kernel")
            for k, v in table.items():
                if k.startswith(key):
                    return int(v)
    return None


def nw_workload(torch, args, dev, hip, family="nw"):
    """Needleman-Wunsch (family "nw") or Saigo-Vert ("sv") affine fwd+bwd on
    the ragged batch of the headline workload (2048 pairs, L1, L2 ~ U[64,
    1024], seed 2, go=-8 ge=-0.5 T=0.5, packed layout).  Roofline: the forward
    streams the scores in and the table out (NW: two planes, M - I and D - I;
    SV: three, M, I - M, D - M), the reverse streams the table and the scores
    in and the map out: 12 / 16 B per cell (NW), 16 / 20 (SV)."""
    from orihime_amd import _lib, ragged
    planes = 2 if family == "nw" else 3
    call = _lib.nw_affine_forward if family == "nw" else _lib.sv_affine_forward
    ws_fn = _lib._ws_nw_affine if family == "nw" else _lib._ws_sv_affine
    params = (-8.0, -0.5, 0.5)
    pairs = ragged_pairs(torch)
    flat, lengths, offsets = ragged.pack(pairs)
    flat, lengths, offsets = flat.to(dev), lengths.to(dev), offsets.to(dev)
    cells = int(sum(p.numel() for p in pairs))
    B, L1max, L2max = len(pairs), 1024, 1024
    ws = torch.empty(int(ws_fn(1, B, L1max, L2max, flat.numel())), dtype=torch.uint8, device=dev)
    value = torch.empty(B, device=dev)
    mp = torch.empty_like(flat)
    g = [torch.empty(B, device=dev) for _ in range(3)]

    class _Nw:
        def step(self):
            call(flat, lengths, offsets, *params, value=value, map=mp,
                                   grad_open=g[0], grad_ext=g[1], grad_temp=g[2], L1max=L1max,
                                   L2max=L2max, workspace=ws)

    launch = _Nw()
    for _ in range(max(1, args.warmup)):
        launch.step()
    torch.cuda.synchronize()
    ms, wall = timed(torch, launch.step, args.steps)
    kd = kernel_durations(torch, _lib, hip, launch, max(3, min(args.steps, 10)))
    fwd_ms, rev_ms = kd["forward"], kd["reverse"]
    gcups = cells / (ms * 1e-3) / 1e9
    byts = {"forward": (4.0 + 4 * planes) * cells, "reverse": (8.0 + 4 * planes) * cells}
    dom = "reverse" if rev_ms >= fwd_ms else "forward"
    dom_ms = rev_ms if dom == "reverse" else fwd_ms
    achieved = byts[dom] / (dom_ms * 1e-3) / 1e9
    log(f"{family} affine fwd+bwd: {ms:.3f} ms/step, {gcups:.1f} GCUPS; forward {fwd_ms:.3f} ms "
        f"({byts['forward'] / fwd_ms / 1e6:.0f} GB/s), reverse {rev_ms:.3f} ms "
        f"({byts['reverse'] / rev_ms / 1e6:.0f} GB/s)")
    line = {
        "metric": f"GCUPS soft-{family.upper()}-affine fwd+bwd (value, map, parameter grads), 1 GPU",
        "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"ragged affine soft-{family.upper()} fwd+bwd, 2048 pairs, L1,L2 ~ "
                               "U[64,1024] (seed 2), gap_open=-8 gap_extend=-0.5 T=0.5, packed",
                   "cells": cells, "wall_ms_per_step": round(wall, 4),
                   "kernel_ms": {k: round(v, 4) for k, v in kd.items()}},
        "roofline": {"bound": "hbm", "kernel": f"{family}_affine_{dom}",
                     "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "traffic": nw_traffic(dom, family),
                     "bytes_per_cell": int(byts[dom] / cells)},
        "cpu_baseline": None,
    }
    # the HVP along an N(0,1) direction (forward-over-reverse pass) on the same batch
    tan = _lib.nw_affine_tangent if family == "nw" else _lib.sv_affine_tangent
    wsf = _lib._ws_nw_affine if family == "nw" else _lib._ws_sv_affine
    wst = torch.empty(int(wsf(2, B, L1max, L2max, flat.numel())), dtype=torch.uint8, device=dev)
    V = torch.randn(flat.shape, generator=torch.Generator().manual_seed(5)).to(dev)
    mt = torch.empty_like(flat)

    def hvp():
        tan(flat, lengths, offsets, *params, direction=V, map_tangent=mt, L1max=L1max,
            L2max=L2max, workspace=wst)

    hvp()
    torch.cuda.synchronize()
    hms, _ = timed(torch, hvp, max(3, args.steps // 2))
    del wst
    log(f"{family} affine HVP: {hms:.3f} ms, {cells / (hms * 1e-3) / 1e9:.1f} GCUPS")
    line["config"]["hvp_ms"] = round(hms, 4)
    if not args.no_cpu:
        log("CPU baseline (about %.0f s)" % args.cpu_seconds)
        line["cpu_baseline"] = cpu_baseline(torch, [p.numpy() for p in pairs], params,
                                            args.cpu_seconds, family)
        log(f"CPU reference: {line['cpu_baseline']['value']} GCUPS on "
            f"{line['cpu_baseline']['cores']} threads")
    print(json.dumps(line), flush=True)


def _grid_time(torch, fn, repeats, target_ms):
    """The reference harness's protocol (benchmarks/grid.py:233-266): five
    warm-up calls, a one-call probe, then `repeats` runs of
    max(3, min(200, ceil(target / probe))) calls each; GPU-event ms per call."""
    import math
    import statistics

    out = None
    for _ in range(5):
        out = fn()
    torch.cuda.synchronize()
    if out is None or not bool(torch.isfinite(out).all()):
        raise RuntimeError("non-finite output")

    def run(n):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(n):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / n

    probe = max(run(1), 0.001)
    n = max(3, min(200, math.ceil(target_ms / probe)))
    samples = [run(n) for _ in range(repeats)]
    return {"iterations": n, "median_ms": statistics.median(samples), "min_ms": min(samples),
            "max_ms": max(samples)}


def grid_workload(torch, args, dev):
    """The reference's Smith-Waterman benchmark grid: every
    rectangular-production-v1 case through the public calls ``sw`` /
    ``sw_affine`` (map surface: forward and reverse sweeps) and ``sw_value``
    / ``sw_affine_value`` (value surface: forward only), without autograd,
    inputs U(-0.5, 0.5) from a CPU generator seeded 0x5A17 + case index
    (benchmarks/grid.py:90-96,174-231).  One JSON line with every result."""
    import orihime_amd as ohm

    rows = []
    for index, (cid, B, L1, L2, explicit) in enumerate(RECT_CASES):
        g = torch.Generator(device="cpu").manual_seed(0x5A17 + index)
        x = (torch.rand((B, L1, L2), generator=g, dtype=torch.float32) - 0.5).to(dev)
        lens = explicit if explicit is not None else [[L1, L2]] * B
        lengths = torch.tensor(lens, dtype=torch.int32, device=dev)
        cells = sum(a * b for a, b in lens)
        for op in ("sw_affine", "sw"):
            for surface in ("value", "map"):
                fn_name = op if surface == "map" else f"{op}_value"
                fn = getattr(ohm, fn_name)
                kw = dict(GRID_PARAMS[op], lengths=lengths)
                with torch.no_grad():
                    t = _grid_time(torch, lambda: fn(x, **kw), args.grid_repeats,
                                   args.grid_target_ms)
                row = {"case": cid, "operator": op, "surface": surface, "batch": B,
                       "cells": cells, **{k: round(v, 5) if isinstance(v, float) else v
                                          for k, v in t.items()},
                       "items_per_second": round(1000.0 * B / t["median_ms"], 1),
                       "gcups": round(cells / t["median_ms"] / 1e6, 3)}
                rows.append(row)
                log(f"{cid} {fn_name}: {t['median_ms']:.4f} ms ({row['gcups']} GCUPS)")
        del x
    line = {"metric": "reference benchmark grid (rectangular-production-v1), median ms per call",
            "unit": "ms", "n_gpus": 1, "higher_is_better": False, "data": "synthetic",
            "dtype": "f32", "results": rows}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload",
                    choices=("ragged", "fixed", "long", "embed", "nw", "sv", "grid"),
                    default="ragged")
    ap.add_argument("--grid-repeats", type=int, default=5)
    ap.add_argument("--grid-target-ms", type=float, default=100.0)
    ap.add_argument("--no-padded", action="store_true", help="skip the padded comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the fixed, long and embedding-scored sub-records")
    ap.add_argument("--no-oracle", action="store_true", help="skip the FP64 oracle errors")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-graph", action="store_true", help="time eager calls instead of a graph")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the automatic launch shape instead of autotuning it on the batch")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 1 GiB device copy first (calibrates HBM byte counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (pairs would be "
              "sharded across GPUs by cost with no exchange but one all-reduce of three "
              "parameter-gradient scalars); run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import orihime_amd  # noqa: F401
    from orihime_amd import _lib, ragged

    if not torch.cuda.is_available():
        sys.exit("bench.py needs a GPU")
    dev = torch.device("cuda:0")
    log(f"device {torch.cuda.get_device_name(0)}")

    if args.copy_probe:
        src = torch.ones(1 << 28, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("copy probe: 1 GiB read + 1 GiB written")

    if args.workload == "embed":
        return embed_workload(torch, args, dev)
    if args.workload in ("nw", "sv"):
        return nw_workload(torch, args, dev, Hip(), args.workload)
    if args.workload == "grid":
        return grid_workload(torch, args, dev)

    pairs = None
    if args.workload == "ragged":
        params = (-8.0, -0.5, 0.5)
        t0 = time.perf_counter()
        pairs = ragged_pairs(torch)
        log(f"generated {len(pairs)} pairs in {time.perf_counter() - t0:.1f} s")
        flat, lengths, offsets = ragged.pack(pairs)
        flat, lengths, offsets = flat.to(dev), lengths.to(dev), offsets.to(dev)
        L1max = L2max = 1024
        cells = int(sum(p.numel() for p in pairs))
        main_launch = Launch(torch, _lib, flat, lengths, offsets, L1max, L2max, params)
        workload = (f"ragged affine soft-SW fwd+bwd, 2048 pairs, L1,L2 ~ U[64,1024] (seed 2), "
                    f"gap_open=-8 gap_extend=-0.5 T=0.5, packed persistent launch")
        cpu_pairs = [p.numpy() for p in pairs]
    elif args.workload == "fixed":
        params = (-11.0, -1.0, 1.0)
        x = fixed_batch(torch)
        cells = x.numel()
        main_launch = Launch(torch, _lib, x.to(dev), None, None, 512, 512, params)
        workload = "fixed-shape affine soft-SW fwd+bwd, [256, 512, 512] (seed 1), go=-11 ge=-1 T=1"
        cpu_pairs = [x[i].numpy() for i in range(x.shape[0])]
    else:
        params = (-11.0, -1.0, 1.0)
        x, _ = long_batch(torch)
        cells = x.numel()
        main_launch = Launch(torch, _lib, x.to(dev), None, None, 4096, 4096, params)
        workload = ("long pairs affine soft-SW fwd+bwd, [16, 4096, 4096] N(0,1) (seed 3), "
                    "go=-11 ge=-1 T=1, checkpointed reverse sweep")
        cpu_pairs = [x[0].numpy()]

    # launch shape: autotuned on this batch (orihime_amd.tune), results do
    # not depend on it
    tuned = _tune(torch, main_launch, params, args)
    log(f"launch shape: {tuned}")

    # warm-up, sanity, graph
    for _ in range(max(1, args.warmup)):
        main_launch.step()
    torch.cuda.synchronize()
    if not bool(torch.isfinite(main_launch.value).all()):
        sys.exit("non-finite values: workspace overflow or invalid input")
    eager = [t.clone() for t in main_launch.outputs()]
    use_graph = not args.no_graph
    if use_graph:
        main_launch.capture()
        for _ in range(2):
            main_launch.graph.replay()
            torch.cuda.synchronize()
            for a, b in zip(eager, main_launch.outputs()):
                if not torch.equal(a, b):
                    sys.exit("graph replay differs from the eager call")
    log("warm-up done")

    # timed region
    run = main_launch.graph.replay if use_graph else main_launch.step
    torch.cuda.synchronize()
    ms, wall_ms = timed(torch, run, args.steps)
    gcups = cells / (ms * 1e-3) / 1e9
    log(f"packed fwd+bwd: {ms:.3f} ms/step ({wall_ms:.3f} wall), {gcups:.1f} GCUPS")
    fwd_ms, _ = timed(torch, main_launch.forward_only, args.steps)
    log(f"forward only: {fwd_ms:.3f} ms, {cells / (fwd_ms * 1e-3) / 1e9:.1f} GCUPS")

    hip = Hip()
    rates = measured_rates(torch, hip)
    log("measured issue rates (T lane-ops/s): " + ", ".join(f"{k} {v}" for k, v in rates.items()))
    hbm_peak = hbm_copy_peak(torch, hip)
    log(f"measured HBM copy bandwidth: {hbm_peak:.0f} GB/s")
    kd = kernel_durations(torch, _lib, hip, main_launch, args.steps)
    log("kernel durations (ms): " + ", ".join(f"{k} {v:.3f}" for k, v in kd.items()))
    # untuned: the library's automatic choice, the fused launch for the
    # per-pair batches (not the long pairs' grid kernels)
    fused = isinstance(tuned, dict) and tuned["teams"] == "fused" or (
        not isinstance(tuned, dict) and os.environ.get("OHM_FUSED", "1") != "0"
        and args.workload != "long")
    if fused:
        # the fused launch runs between the plan and forward marks
        kd["fused"] = kd["forward"]
        dom = "fused"
    else:
        dom = max(("forward", "reverse"), key=lambda k: kd[k])
    t_dom = kd[dom] * 1e-3
    trans_peak = rates["exp"] * 1e12 if rates else TRANS_PEAK
    sweeps = ("forward", "reverse")  # one step = one forward and one reverse sweep per cell
    step_trans = sum(TRANS_PER_CELL[k] for k in sweeps)
    step_bytes = sum(BYTES_PER_CELL[k] for k in sweeps)
    step_roof = min(hbm_peak * 1e9 / step_bytes, trans_peak / step_trans) / 1e9
    wl_name = "ragged2048" if args.workload == "ragged" else args.workload
    shape = (tuned["waves"], tuned["checkpoint_stride"]) if isinstance(tuned, dict) else None
    traffic, traffic_src = profiled_traffic(wl_name, dom, shape)
    valu, valu_src = profiled_valu(wl_name, dom, shape)
    issue_peak = max(rates["fma"], rates["dpp"]) * 1e12 / 64 if rates else VALU_LANE_OPS / 64
    step_achieved = step_trans * cells / (ms * 1e-3)
    dom_hbm = BYTES_PER_CELL[dom] * cells / t_dom / 1e9
    roofline = {
        # binding roof of the step: the reference's transcendentals per cell
        # (22 for a forward and a reverse sweep) against the measured exp rate
        "bound": "transcendental",
        "achieved": round(step_achieved / 1e12, 4),
        "peak": round(trans_peak / 1e12, 4),
        "unit": "Top/s",
        "frac": round(step_achieved / trans_peak, 4),
        "traffic": traffic,
        "per_cell": step_trans,
        "peak_source": "measured v_exp_f32 issue rate" if rates else "quarter-rate assumption",
        "scope": "whole step (graph replay), " + ("one fused forward+reverse launch" if fused
                                                 else "forward and reverse launches"),
        "step_roof_gcups": round(step_roof, 1),
        "step_frac": round(gcups / step_roof, 4),
        "hbm": {
            "bound": "hbm",
            "kernel": KERNEL_NAMES[dom],
            "kernel_ms": round(kd[dom], 4),
            "algorithmic_bytes": BYTES_PER_CELL[dom] * cells,
            "achieved": round(dom_hbm, 2),
            "peak": round(hbm_peak, 1),
            "peak_source": "measured device-to-device copy (bytes read + written)",
            "peak_spec": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom_hbm / hbm_peak, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "traffic_frac": None if traffic is None else round(traffic / t_dom / (hbm_peak * 1e9), 4),
        },
        "dominant_kernel": {
            "kernel": KERNEL_NAMES[dom],
            "kernel_ms": round(kd[dom], 4),
            "transcendental_per_cell": TRANS_PER_CELL[dom],
            "transcendental_frac": round(TRANS_PER_CELL[dom] * cells / t_dom / trans_peak, 4),
        },
        # what binds these kernels: VALU issue (profiled instruction count of
        # the dominant kernel against the measured full-rate issue)
        "issue": None if valu is None else {
            "bound": "valu",
            "instructions": valu,
            "instructions_source": valu_src,
            "achieved": round(valu / t_dom / 1e9, 2),
            "peak": round(issue_peak / 1e9, 2),
            "unit": "G wave-instructions/s",
            "frac": round(valu / t_dom / issue_peak, 4),
        },
        "measured_issue_rates_tops": rates,
    }

    extra = {}
    if pairs is not None and not args.no_oracle:
        t0 = time.perf_counter()
        maps = [m.cpu().numpy() for m in ragged.unpack(main_launch.map, lengths, offsets)]
        extra["oracle_err"] = oracle_errors(cpu_pairs, (main_launch.value.cpu().numpy(), maps), params)
        log(f"FP64 oracle errors ({time.perf_counter() - t0:.1f} s): {extra['oracle_err']}")
        del maps
    if args.workload == "ragged" and not args.no_padded:
        # the same batch padded to [B, 1024, 1024], cells beyond the lengths masked
        from orihime_amd.operator import _normalize_masked

        B = len(pairs)
        pad = torch.full((B, L1max, L2max), float("-inf"), device=dev)
        for i, p in enumerate(ragged.unpack(flat, lengths, offsets)):
            pad[i, :p.shape[0], :p.shape[1]] = p
        pad = _normalize_masked(pad, params[2]).contiguous()
        padded = Launch(torch, _lib, pad, None, None, L1max, L2max, params)
        extra["padded_launch_config"] = _tune(torch, padded, params, args)
        for _ in range(max(1, args.warmup)):
            padded.step()
        torch.cuda.synchronize()
        if use_graph:
            padded.capture()
        prun = padded.graph.replay if use_graph else padded.step
        pms, _ = timed(torch, prun, max(3, args.steps // 4))
        pg = cells / (pms * 1e-3) / 1e9
        log(f"padded+mask fwd+bwd: {pms:.3f} ms/step, {pg:.1f} GCUPS (useful cells)")
        vdiff = float(((padded.value - main_launch.value).abs()
                       / main_launch.value.abs().clamp_min(1e-3)).max())
        extra.update({"padded_ms_per_step": round(pms, 4), "padded_gcups": round(pg, 2),
                      "ragged_over_padded": round(gcups / pg, 3),
                      "padded_vs_packed_value_rel_diff": vdiff})
        if not args.no_oracle:
            idx = [0, 1000, B - 1]
            extra["padded_oracle_err"] = _padded_oracle_errors(
                pad[idx].cpu().numpy(), (padded.value.cpu().numpy()[idx], padded.map[idx].cpu().numpy()),
                params, list(range(3)))
            extra["padded_oracle_err"]["pairs"] = idx
        del padded, pad
        torch.cuda.empty_cache()

    if args.workload == "long":
        extra["long"] = long_record(torch, _lib, hip, args, dev)
    if args.workload == "ragged" and not args.no_configs:
        del main_launch.graph
        extra["configs"] = {}
        for name, fn in (("fixed_256x512", lambda: fixed_record(torch, _lib, hip, args, dev)),
                         ("long_16x4096", lambda: long_record(torch, _lib, hip, args, dev)),
                         ("embed_512", lambda: embed_record(torch, hip, args, dev))):
            t0 = time.perf_counter()
            extra["configs"][name] = fn()
            log(f"{name} done in {time.perf_counter() - t0:.1f} s")
            torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline (about {args.cpu_seconds:.0f} s on {host_threads()} threads, then 1)")
        cpu = cpu_baseline(torch, cpu_pairs, params, args.cpu_seconds)
        log(f"CPU {cpu['kind']}: {cpu['value']} GCUPS on {cpu['threads']} threads, "
            f"{cpu['single_thread']['value']} on 1 ({cpu['cpu_model']})")

    line = {
        "metric": METRIC,
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "cells": cells,
            "forward_gcups": round(cells / (fwd_ms * 1e-3) / 1e9, 3),
            "kernel_ms": {k: round(v, 4) for k, v in kd.items()},
            "hip_graph": use_graph,
            "launch_config": tuned,
            "wall_ms_per_step": round(wall_ms, 4),
            **extra,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
