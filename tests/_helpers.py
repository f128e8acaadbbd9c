"""Shared fixtures-by-function for the test modules."""

import functools
import os

import numpy as np

from create_rate_subs_amd import workloads
from create_rate_subs_amd.create_rate_subs import load_mechanism

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@functools.lru_cache(maxsize=None)
def mechanism(name):
    """(specs, reacs, fwd_spec_mapping) of a synthetic mechanism."""
    mpath, tpath = workloads.mechanism_files(name)
    _e, specs, reacs, fwd = load_mechanism(
        mpath, tpath, last_spec=workloads.last_species(name))
    return specs, reacs, fwd


@functools.lru_cache(maxsize=None)
def oracle(name):
    from oracle.build import Oracle, lib_path, build
    if not os.path.exists(lib_path(name)):
        build(name)
    return Oracle(name)


def hip_lib(name):
    from create_rate_subs_amd import build, runtime
    return runtime.load(build.lib_path(name))


def to_dev(*arrays):
    import torch
    return [torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                            device='cuda') for a in arrays]


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()
