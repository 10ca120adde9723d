"""Work accounting of bench.py's roofline (no GPU): the algorithmic flops
of a stepper launch follow from its label, counting every product once."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope='module')
def bench():
    spec = importlib.util.spec_from_file_location('bench_module', os.path.join(ROOT, 'bench.py'))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_chunked_launch_counts_each_product_once(bench):
    # 8 chunks of 15 products each (16 saves per chunk, one step per save)
    # plus the 7 products of the chain of chunk starts: 127 products, the
    # same as the plain 127-step trajectory (saves at 0 .. 127 steps)
    M = 147
    got = bench.launch_flops('chunks prop mixed M=147 nc=8 B=1 steps=15 nsave=16')
    assert got == 8.0 * M * M * 127
    got = bench.launch_flops('chunks prop mixed M=147 nc=11 B=1 steps=11 nsave=12')
    assert got == 8.0 * M * M * (11 * 11 + 10)


def test_propagate_launch_counts(bench):
    # the chain of chunk starts: one right-hand side, 7 propagator products
    assert bench.launch_flops('propagate prop fp64 M=147 R=1 B=1 steps=7 nsave=8') == \
        8.0 * 147 * 147 * 7
    # classical RK4: four products per step, every right-hand side and problem
    assert bench.launch_flops('propagate rk4 fp64 M=7 R=3 B=2 steps=10 nsave=11') == \
        8.0 * 7 * 7 * 4 * 10 * 3 * 2


def test_traffic_lookup_by_launch_shape(bench):
    assert bench.launch_traffic('chunks prop mixed M=147 nc=8 B=1 steps=15 nsave=16') == \
        bench.PMC_CHUNKS[('mixed', 147, 8, 16)]
    assert bench.launch_traffic('propagate prop fp64 M=147 R=1 B=1 steps=7 nsave=8') == \
        bench.PMC_PROPAGATE[('fp64', 147, 1, 8)]
    assert bench.launch_traffic('chunks prop mixed M=31 nc=8 B=1 steps=15 nsave=16') is None
