"""Helpers of the GPU tests (inputs come from the oracle, on the host)."""
import sys

import numpy as np
import torch


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()
