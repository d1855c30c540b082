"""Train the learned seeder on the GPU and write its weights.

    python semper-dic-engine_amd/train_seeder.py [--steps 8000] [--batch 256]
        [--lr 1e-3] [--seed 20240] [--out semper-dic-engine_amd/seeder_weights.bin]

The pairs are the engine's own: a 2048 x 2048 reference of its synthetic
speckle (the workloads' statistics, noise 4 grey levels), cut into patches at
pyramid level 2 and warped through a Keys bicubic interpolator by random
affine maps (shift up to 28 px per component, displacement gradients up to
0.12), on half of the pairs plus a smooth heterogeneous field (a Gaussian
bump per displacement component, up to 8 px, 100-400 px wide, centred
within one width of the patch), with fresh noise on the deformed samples.
The target is the field's first-order warp at the patch centre. 8000 steps
of 256 are 2.05 M pairs. The forward and backward passes, the loss and Adam are the
library's own HIP kernels (csrc/seeder.hip). Prints the mean loss of every
100 steps.
"""
import argparse
import ctypes
import importlib.util
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20240)
    ap.add_argument("--out", default=os.path.join(HERE, "seeder_weights.bin"))
    args = ap.parse_args()
    spec = importlib.util.spec_from_file_location("semper_amd", os.path.join(HERE, "semper_amd.py"))
    sa = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(sa)
    dev = sa.load_dev()
    if dev.sdk_device_count() < 1:
        sys.exit("no GPU visible")
    cap = args.steps // 100 + 1
    losses = (ctypes.c_float * cap)()
    t0 = time.perf_counter()
    k = dev.sdk_seeder_train(args.seed, args.steps, args.batch, ctypes.c_float(args.lr), args.out.encode(), losses,
                             cap)
    dt = time.perf_counter() - t0
    if k < 0:
        sys.exit(f"training failed ({k})")
    for i in range(k):
        print(f"steps {min((i + 1) * 100, args.steps):6d}  loss {losses[i]:.5f}", flush=True)
    print(f"trained {args.steps} x {args.batch} pairs in {dt:.1f} s -> {args.out}", flush=True)


if __name__ == "__main__":
    main()
