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


def test_traffic_lookup_by_launch_label(bench):
    label = 'propagate powers M=49 B=1 R=7 e=0,100,200,500'
    assert bench.launch_traffic(label) == bench.PMC_TRAFFIC[label]
    assert bench.launch_traffic('chunks prop mixed M=31 nc=8 B=1 steps=15 nsave=16') is None


def test_operator_program_counts(bench):
    # Horner's three products, q = 2 (one squaring), c = 16 (four squarings)
    assert bench.launch_flops('small propagator M=7 B=1 q=2 c=16') == 8.0 * 7 ** 3 * (3 + 1 + 4)
    # c = 12 = 8 + 4: three squarings and one product
    assert bench.launch_flops('small propagator M=7 B=2 q=1 c=12') == \
        8.0 * 7 ** 3 * 2 * (3 + 0 + 4)
    # Horner, squarings up to P^256 (8), and one state product per set bit
    # of 100 (3), 200 (3) and 500 (6), seven states
    got = bench.launch_flops('propagate powers M=49 B=1 R=7 e=0,100,200,500')
    assert got == 8.0 * (49 ** 3 * (3 + 8) + 49 * 49 * 7 * 12)
    assert bench.launch_flops('zgemm fp64 m=255 n=255 k=154 batch=4 trans=2') == \
        8.0 * 255 * 255 * 154 * 4


def test_kernel_bounds(bench):
    assert bench.kernel_bound('zgemm fp64 m=147 n=147 k=147 batch=256') == ('mfma', 78.6)
    assert bench.kernel_bound('zgemm fp64 m=147 n=147 k=147 batch=1') == ('valu', 78.6)
    assert bench.kernel_bound('chunks prop mixed M=147 nc=8 B=1 steps=15 nsave=16') == \
        ('valu', 157.3)
    assert bench.kernel_bound('propagate powers M=49 B=1 R=7 e=0,100') == ('valu', 78.6)
