"""Bounds-checked CPU run of every libpdx kernel body (AddressSanitizer + UBSan).

The per-particle functions that the HIP kernels wrap (simul_amd/csrc/pdx_device.h)
and the library's host layout (pdx_layout.h) are compiled with g++ and swept
over every particle of the cases below, with each array allocated at exactly the
size the library allocates; any out-of-bounds access fails the run.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import lattice
from simul_amd import workloads as W

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_check")
sys.path.insert(0, HERE)
from inputs import write_input  # noqa: E402

EXE = os.path.join(HERE, "pdx_cpu_check")


@pytest.fixture(scope="module")
def exe():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return EXE


def _run(exe, tmp_path, name, **kw):
    path = str(tmp_path / f"{name}.bin")
    write_input(path, **kw)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, path, "3"], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ok n=" in r.stdout
    return r.stdout


def _inputs(X, E=1000.0, nu=0.3, horizon=3.015, dx=1.0, fixed=None, fext=None, u_amp=1e-3, seed=0):
    n, dim = X.shape
    Kb, mu = W.bulk_shear(np.full(n, E), nu)
    fx = np.zeros(n, np.uint8) if fixed is None else fixed
    u0 = np.random.default_rng(seed).uniform(-u_amp, u_amp, X.shape)
    return dict(X=X, vol=np.full(n, dx ** dim), mass=np.full(n, dx ** dim), bulk=Kb, shear=mu,
                cst=18 * Kb / (np.pi * horizon ** 4), fixed=fx, fext=fext, u0=u0, horizon=horizon)


@pytest.mark.parametrize("model,form", [("linear_elastic", 0), ("neo_hookean", 0), ("linear_elastic", 1)])
def test_kernels_in_bounds_lattice(exe, tmp_path, model, form):
    X = lattice(11, 9, 8, jitter=0.2, seed=2)
    fx = np.zeros(len(X), np.uint8)
    fx[X[:, 2] < 0.5] = 7
    fx[:5] = 3
    fe = np.zeros_like(X)
    fe[X[:, 2] > 6.5, 2] = -1.0
    _run(exe, tmp_path, f"lat_{model}_{form}", model=model, form=form, **_inputs(X, fixed=fx, fext=fe))


def test_kernels_in_bounds_j2(exe, tmp_path):
    X = lattice(8, 8, 8, jitter=0.1, seed=3)
    n = len(X)
    _run(exe, tmp_path, "j2", model="j2_plasticity", sy=np.full(n, 0.5), hiso=np.full(n, 50.0),
         hkin=np.full(n, 10.0), **_inputs(X, u_amp=5e-3))


def test_kernels_in_bounds_batch_and_isolated(exe, tmp_path):
    parts = [lattice(5, 5, 5), lattice(9, 4, 3) + 100.0, np.array([[500.0, 0, 0]]), lattice(7, 6, 5, jitter=0.2)]
    off = np.cumsum([0] + [len(p) for p in parts])
    X = np.concatenate(parts)
    _run(exe, tmp_path, "batch", model="neo_hookean", offsets=off, **_inputs(X))


def test_kernels_in_bounds_2d(exe, tmp_path):
    g = np.arange(17.0)
    X = np.stack(np.meshgrid(g, g[:9], indexing="ij"), -1).reshape(-1, 2)
    _run(exe, tmp_path, "two_d", model="neo_hookean", **_inputs(X))


def test_kernels_in_bounds_vertebra_coarse(exe, tmp_path):
    spec = W.vertebra(dx=1.0)
    inp = _inputs(spec.X, horizon=spec.horizon, dx=spec.dx, fixed=spec.fixed, fext=spec.fext)
    inp["bulk"], inp["shear"] = spec.bulk, spec.shear
    _run(exe, tmp_path, "vertebra", model="neo_hookean", **inp)


def test_layout_bank_colours_and_lattice_quantisation(exe):
    """The LDS slot colours of the halo kernels: any shifted 2x2x4 (2-D: 4x4)
    box of cells holds 16 colours, the Morton quantisation at the median
    spacing puts each aligned run of 16 keys of a jittered lattice on one box,
    and every SELL slice width is its longest row rounded up to a multiple of
    4 slots (tests/cpu_check/layout_check.cpp)."""
    r = subprocess.run([os.path.join(HERE, "layout_check")], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert r.returncode == 0, r.stderr[-2000:]
    assert "layout ok" in r.stdout


def test_halo_tables_in_bounds(exe):
    """The halo tables of the LDS-staged kernels from the library's host
    builder on lattices, a padded three-case batch, a 2-D plate, a coarse
    vertebral body and a wide horizon, with broken bonds: every table entry
    within its block's LDS slots, and the kernels' row loop (every bond group
    and split, with its look-ahead loads) inside the bk array the library
    allocates (tests/cpu_check/halo_check.cpp, AddressSanitizer / UBSan)."""
    r = subprocess.run([os.path.join(HERE, "halo_check")], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1"))
    assert r.returncode == 0, r.stderr[-2000:]
    assert "halo ok" in r.stdout
