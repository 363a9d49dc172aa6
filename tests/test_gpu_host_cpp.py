"""The C++ host layer on the device: an H2 stochastic-reconfiguration run
through GpuSampler + StochasticReconfiguration + VmcRunner (C++) gives the
same parameters and energies, bit for bit, as the same schedule driven
from Python over the same C ABI; GpuWaveFunction and GpuSlaterJastrow
evaluations equal mole_eval's."""
import json
import os
import subprocess

import numpy as np
import pytest

import mole_amd as M

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(__file__), "..", "mole_amd", "host", "_build", "host_tests")


def test_cpp_vmc_runner_matches_python_schedule():
    iters, W, sweeps, block, box, step, eps, seed = 4, 2048, 60, 10, 0.5, 0.05, 1e-3, 3
    r = subprocess.run([BIN, "--gpu", str(iters), str(W), str(sweeps), str(block), str(box), str(step), str(eps),
                        str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    p = np.array([1.0, 0.5, 0.3])
    energies = []
    with M.Context([M.h2(W, *p)], seed=seed, step=box, block_size=block) as ctx:
        ctx.init_walkers()
        for _ in range(iters):
            ctx.reset_stats()
            ctx.run(sweeps, block)
            st = ctx.stats()
            energies.append(st["sum_e"] / st["n_samples"])
            dp, _ = M.parameter_update(ctx.sr_sums_raw(), M.OPT_SR, step, M.SR_TEXTBOOK, eps)
            p = p + dp
            ctx.set_params(p)
    assert np.array_equal(np.array(out["params"]), p)
    assert np.array_equal(np.array(out["energies"]), np.array(energies))
    # GpuWaveFunction (traits) against mole_eval through Python
    x = np.array([[0.3, -0.2, 0.5], [-0.4, 0.1, -0.6]])
    with M.Context([M.helium(1)], seed=0, step=0.5, block_size=1) as ctx:
        e = ctx.eval(x.reshape(1, 6))
    assert out["he_value"] == e["value"][0]
    assert out["he_laplacian"] == e["lap"][0]
    assert out["he_local_energy"] == e["e_local"][0]
    assert np.array_equal(np.array(out["he_grad"]), e["grad"].reshape(-1))
    # GpuSlaterJastrow (traits) against mole_eval of the same Slater-Jastrow
    # system through Python, and a copy evaluating on its own context
    from mole_amd import systems as S
    hs = S.SJSystem([(0.0, 0.0, 0.0)], [2], 1, 1, [1.6875], [0], [[1.0]], (0.5, 0.0, 0.3, 0.0, 1.0))
    with M.Context([M.slater_jastrow(hs, 1)], seed=0, step=0.5, block_size=1) as ctx:
        e = ctx.eval(x.reshape(1, 6))
    assert out["sj_value"] == e["value"][0] and out["sj_laplacian"] == e["lap"][0]
    assert out["sj_local_energy"] == e["e_local"][0]
    assert np.array_equal(np.array(out["sj_grad"]), e["grad"].reshape(-1))
    assert np.array_equal(np.array(out["sj_pgrad"]), e["pgrad"].reshape(-1))
    assert out["sj_copy_same"] == 1
    # the streaming Runner::run of 6 blocks: Log::log after each of the 5
    # sampled blocks, the last call seeing all 5 block means
    assert out["stream_calls"] == 5 and out["stream_blocks"] == 5 and out["result_blocks"] == 5
    # examples/custom_operator.rs through GpuSampler<GpuSlaterJastrow> with a
    # device HarmonicHamiltonian; DmcRunner on hydrogen; the panic path
    assert abs(out["sho_energy"] - 1.5) < 1e-12
    assert abs(out["dmc_energy"] + 0.5) < 1e-12 and out["dmc_blocks"] == 3
    assert out["dmc_panics"] == 1
