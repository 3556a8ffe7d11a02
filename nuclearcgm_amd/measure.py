"""Measurement drivers for profiler runs (rocprofv3 kernel traces and
hardware counters):

    python -m nuclearcgm_amd.measure solve [--iters N] [--wstep-iters K] [--no-train]
        trains the learned preconditioner (as bench.py does), then runs the
        accelerated Ca-40 40^3 solve for N HF iterations and prints the live
        HIP-event averages of its stencil and block-product launches;
        --no-train: Jacobi-PCG W-block, no training (few dispatches, for
        hardware-counter passes)
    python -m nuclearcgm_amd.measure copy [--mib M]
        one device-to-device copy of M MiB (a kernel of known traffic, the
        calibration of the FETCH_SIZE / WRITE_SIZE counters)
    python -m nuclearcgm_amd.measure stencil [--cols M] [--reps R]
        R launches of the H.psi stencil on M device-resident FP32 columns of
        the Ca-40 lattice
    python -m nuclearcgm_amd.measure gram [--reps R]
        R block products of device-resident FP32 state blocks of the Ca-40
        lattice for the shapes of the GCG pass (k x k Hermitian, k x k, ...)
"""
import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["solve", "copy", "stencil", "gram"])
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--wstep-iters", type=int, default=1)
    ap.add_argument("--no-train", action="store_true")
    ap.add_argument("--mib", type=int, default=512)
    ap.add_argument("--cols", type=int, default=32)
    ap.add_argument("--reps", type=int, default=50)
    a = ap.parse_args(argv)
    if a.what == "copy":
        import torch

        x = torch.ones(a.mib * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(f"copied {a.mib} MiB", flush=True)
        return
    sys.path.insert(0, ROOT)
    import bench
    from nuclearcgm_amd import config as C
    from nuclearcgm_amd import hip

    ctx = hip.Context(0)
    if a.what == "stencil":
        cfg = C.workload("ca40", precision=C.PREC_MIXED)
        p = ctx.problem(cfg)
        n3 = cfg.n ** 3
        import numpy as np

        rng = np.random.default_rng(0)
        f = 20.0 + rng.uniform(0, 1, n3)
        gf = rng.uniform(-1, 1, 3 * n3)
        U = rng.uniform(-50, 0, n3)
        W = rng.uniform(-1, 1, 3 * n3)
        for q in (0, 1):
            p.set_fields(q, f, gf, U, W)
        ms = p.bench_hpsi(0, a.cols, C.PREC_MIXED, a.reps)
        print(f"stencil {a.cols} columns: {ms * 1e3:.2f} us per launch", flush=True)
        return
    if a.what == "gram":
        cfg = C.workload("ca40", precision=C.PREC_MIXED)
        p = ctx.problem(cfg)
        M = cfg.M_n
        n3 = cfg.n ** 3
        for ka, kb, low in ((3 * M, 3 * M, True), (3 * M, 3 * M, False), (M, M, True), (3 * M, M, False),
                            (2 * M, 2 * M, True)):
            ms = p.bench_gram(0, ka, kb, low, a.reps)
            mb = (ka + (0 if low else kb)) * n3 * 16 / 1e6
            print(f"gram {ka}x{kb} {'hermitian' if low else 'general'}: {ms * 1e3:.2f} us per product, "
                  f"{mb:.1f} MB algorithmic, {mb / ms / 1e3:.0f} GB/s", flush=True)
        return
    if a.no_train:
        cfg = bench.accelerated_config("ca40", C.W_JACOBI_CG, 0, C.MIX_BROYDEN, 8)
    else:
        setup = bench.train_preconditioner(ctx)
        print(f"trained: loss {setup['loss_first']:.3f} -> {setup['loss_last']:.3f}", flush=True)
        cfg = bench.accelerated_config("ca40", iters=a.wstep_iters)
    cfg.max_iter = a.iters
    p = ctx.problem(cfg)
    p.kernel_stats(enable=1)
    res = p.scf_run(*bench.make_guess(cfg))
    ks = p.kernel_stats(enable=0)
    print(f"solve: {res[4].iterations} iterations, E {res[4].terms[C.E_INT]:.6f}", flush=True)
    for k, v in ks.items():
        n = max(v["launches"], 1)
        print(f"live {k}: {v['launches']} launches, {v['ms'] * 1e3 / n:.2f} us avg, "
              f"{v['bytes'] / n / 1e6:.2f} MB algorithmic per launch, "
              f"{v['bytes'] / max(v['ms'], 1e-12) / 1e6:.1f} GB/s", flush=True)


if __name__ == "__main__":
    main()
