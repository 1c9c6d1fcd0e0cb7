"""Benchmark: 2919-feature vectors/s of the MI355X engine.

Workload: the long-set configuration -- 256x256 uint16 synthetic images (4
classes x 100, grating periods 6/12/24/48 px, seed 2002), the full 2919
feature DAG (11 transform nodes, 124 feature groups) computed in FP64.  One
step = one pass of the DAG over one batch of images already resident in
HBM (wndgpu_compute_device); value = images/s over the timed steps.

Also reported: the roofline position of the dominant kernel (per-launch HIP
events on the engine's stream during the timed region) and the CPU baseline
(the FP64 C++ oracle, one image per thread on the host cores, on a bounded
sample of the same images).
"""
import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'wnd-charm_amd'))


def cpu_baseline(imgs, seconds=12.0):
    """Oracle (CPU restatement) throughput on host cores; a bounded sample."""
    lib = ctypes.CDLL(os.path.join(ROOT, 'oracle', 'liboracle.so'))
    lib.orc_features.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    cores = min(os.cpu_count() or 1, 16)
    done = [0]
    lock = threading.Lock()
    stop = time.time() + seconds
    planes = [np.ascontiguousarray(i.astype(np.float64)) for i in imgs]

    def worker(k):
        out = np.zeros(2919)
        j = k
        while time.time() < stop:
            p = planes[j % len(planes)]
            lib.orc_features(1, ctypes.c_void_p(p.ctypes.data), p.shape[0], p.shape[1], ctypes.c_void_p(out.ctypes.data))
            with lock:
                done[0] += 1
            j += cores

    t0 = time.time()
    th = [threading.Thread(target=worker, args=(k,)) for k in range(cores)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    el = time.time() - t0
    return {'value': done[0] / el, 'unit': 'feature vectors/s', 'cores': cores, 'kind': 'port',
            'sample': '%d long-set vectors of 256x256 images in %.1f s, one image per thread' % (done[0], el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--batch', type=int, default=400)
    ap.add_argument('--no-cpu', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: every run here uses a single MI355X; '
              'multi-GPU sharding is described in DESIGN.md but not run', file=sys.stderr)
        sys.exit(2)
    import torch
    import wndcharm
    from wndcharm import synthetic

    imgs, cls = synthetic.config_images('long_256', limit=args.batch)
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSetLong()
    nf = plan.n_features
    dev_imgs = [torch.from_numpy(im.view(np.int16)).cuda() for im in imgs]
    out = torch.zeros((len(imgs), nf), dtype=torch.float64, device='cuda')
    structs = (wndcharm._Image * len(imgs))()
    for i, t in enumerate(dev_imgs):
        structs[i] = wndcharm._Image(ctypes.c_void_p(t.data_ptr()), t.shape[0], t.shape[1], t.shape[1], 0)
    lib = wndcharm._load()
    ctx = wndcharm._context()
    status = np.zeros(len(imgs), np.int32)

    def step():
        rc = lib.wndgpu_compute_device(ctx, plan.handle, structs, len(imgs), ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(status.ctypes.data))
        if rc != 0:
            raise RuntimeError(lib.wndgpu_last_error(ctx).decode())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_step = el * 1000 / args.steps
    value = len(imgs) * args.steps / el
    # per-launch times of one more pass (events on the engine stream)
    os.environ['WNDGPU_TIMING'] = '1'
    step()
    times = wndcharm.last_node_times()
    del os.environ['WNDGPU_TIMING']
    agg = {}
    for name, ms in times:
        grp = name.split(' (')[0]
        agg[grp] = agg.get(grp, 0.0) + ms
    top = sorted(agg.items(), key=lambda kv: -kv[1])
    line = {
        'metric': 'feature vectors/sec (2919-feat)', 'value': round(value, 3), 'unit': 'feature vectors/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(ms_step, 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic (seeded grating/blob/Poisson generator, 256x256 uint16)',
        'config': {'workload': 'long-set 2919 features, %d images 256x256 per step, plain FP64' % len(imgs),
                   'images_per_step': len(imgs), 'set': 'long', 'precision': 'fp64'},
        'stage_ms': {k: round(v, 3) for k, v in top[:12]},
    }
    if not args.no_cpu:
        line['cpu_baseline'] = cpu_baseline(imgs[:64])
    print(json.dumps(line))


if __name__ == '__main__':
    main()
