"""The reference benchmark's 102 workloads (fp32, alpha=2.3, beta=4.2, 50-60
million elements each; test/benchmark/inc/hptt/benchmark/benchmark_trans.h:
24-140, test/benchmark/inc/hptt/perf-test/test_perf_util.h:20-22) at full size
on the GPU, bit for bit against the oracle."""
import numpy as np
import pytest

import oracle
from gpu_util import run_plan
from hptt_amd.workloads import (REFERENCE_SWEEP_ALPHA, REFERENCE_SWEEP_BETA,
                                reference_sweep)

pytestmark = pytest.mark.gpu

CASES = reference_sweep()


@pytest.mark.parametrize("name,size,perm", CASES, ids=[c[0] for c in CASES])
def test_reference_workload(device, name, size, perm):
    n = int(np.prod(size))
    seed = 31 * CASES.index((name, size, perm))
    A = oracle.uniform(n, np.float32, 7000 + seed, nthreads=16)
    B0 = oracle.uniform(n, np.float32, 9000 + seed, nthreads=16)
    exp = B0.copy()
    assert oracle.transpose(A, exp, size, perm, REFERENCE_SWEEP_ALPHA, REFERENCE_SWEEP_BETA,
                            nthreads=16) == 0
    got, info = run_plan(A, B0, size, perm, REFERENCE_SWEEP_ALPHA, REFERENCE_SWEEP_BETA)
    bad = oracle.first_mismatch(got, exp)
    assert bad == -1, (name, size, perm, bad)
