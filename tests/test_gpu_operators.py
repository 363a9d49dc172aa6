"""Local operators on the device (mole_set_hamiltonian, mole_set_observables,
mole_read_observables) and Gaussian basis functions, against the oracle
(orc_sj_operator, the parity runner with a custom Hamiltonian and
observables): accept decisions bit for bit, local values to 1e-10; the
oscillator of examples/custom_operator.rs at zero variance in VMC and DMC."""
import numpy as np
import pytest

import mole_amd as M
from mole_amd import systems as S
from oracle import sj_ffi as SJ

pytestmark = pytest.mark.gpu


def gto_atom():
    return S.SJSystem([(0.0, 0.0, 0.0)], [2], 1, 1, [0.8, 1.7, 2.3], [0, 0, 0], [[0.6, 0.3, 0.5]],
                      (0.5, 0.25, 1.0, 0.3, 1.0), name="He GTO", kind=[1, 0, 1])


def gto_molecule():
    return S.SJSystem([(0.0, 0.0, -0.7), (0.0, 0.0, 0.7)], [2, 1], 2, 1, [0.5, 1.2, 0.7, 1.5], [0, 0, 1, 1],
                      [[0.7, 0.2, 0.4, 0.1], [0.1, -0.6, 0.3, 0.8]], (0.5, 0.25, 1.0, 0.3, 1.0),
                      name="GTO molecule", kind=[1, 1, 1, 1])


def device_run(sj, W, steps, B, *, h=None, obs=(), proposal=M.PROPOSAL_BOX, side=0.6, lu=16, seed=0,
               precision=M.PREC_FP64):
    with M.Context([M.slater_jastrow(sj, W)], seed=seed, step=side, block_size=B, flags=M.FLAG_TRACE,
                   record_sweeps=steps, lu_period=lu, proposal=proposal, precision=precision) as ctx:
        if h is not None:
            ctx.set_hamiltonian(h)
        if obs:
            ctx.set_observables(list(obs))
        ctx.init_walkers()
        ctx.run(steps)
        masks, el = ctx.trace(steps, steps - B)
        o = ctx.read_observables() if obs else None
        return dict(masks=masks, el=el, stats=ctx.stats(), obs=o)


def oracle_run(sj, W, steps, B, *, h=None, obs=(), proposal=SJ.BOX, side=0.6, lu=16, seed=0):
    return SJ.run(SJ.SjWave(sj), seed=seed, walkers=W, steps=steps, block_size=B, box_side=side, lu_period=lu,
                  n_threads=8, proposal=proposal, tau=side, hamiltonian=h, observables=obs)


@pytest.mark.parametrize("make,proposal", [(gto_atom, M.PROPOSAL_BOX), (gto_molecule, M.PROPOSAL_BOX),
                                           (gto_molecule, M.PROPOSAL_DIFFUSE)])
def test_gto_parity_with_oracle(make, proposal):
    sj = make()
    side = 0.6 if proposal == M.PROPOSAL_BOX else 0.05
    print(f"gto parity {sj.name} proposal={proposal}", flush=True)
    d = device_run(sj, 48, 30, 10, proposal=proposal, side=side)
    o = oracle_run(sj, 48, 30, 10, proposal=SJ.BOX if proposal == M.PROPOSAL_BOX else SJ.DIFFUSE, side=side)
    assert np.array_equal(d["masks"], o["accept_mask"])
    assert np.allclose(d["el"], o["e_local"], rtol=1e-10, atol=1e-10)


def test_custom_electronic_hamiltonian_matches_built_in():
    sj = S.atom(4)
    a = device_run(sj, 64, 30, 10)
    b = device_run(sj, 64, 30, 10, h=M.electronic_hamiltonian())
    assert np.array_equal(a["masks"], b["masks"])
    assert np.allclose(a["el"], b["el"], rtol=1e-12, atol=1e-12)


def test_observables_against_oracle_and_energy():
    sj = S.atom(3)
    obs = [[(M.TERM_KINETIC, 1.0)], [(M.TERM_IONIC_POTENTIAL, 1.0)], [(M.TERM_ELECTRONIC_POTENTIAL, 1.0)],
           [(M.TERM_HARMONIC, 1.0, 1.0, 0.0, 0.0, 0.0)], [(M.TERM_RADIAL_POWER, 1.0, -1, 0.0, 0.0, 0.0)],
           [(M.TERM_DIPOLE, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)]]
    d = device_run(sj, 64, 40, 10, obs=obs)
    o = oracle_run(sj, 64, 40, 10, obs=obs)
    s1, s2, n = d["obs"]
    assert n == o["n_samples"] == d["stats"]["n_samples"]
    assert np.allclose(s1, o["obs_sum"], rtol=1e-10, atol=1e-9)
    assert np.allclose(s2, o["obs_sum2"], rtol=1e-10, atol=1e-9)
    # kinetic + ionic + electronic = the sampled energy
    assert abs(s1[:3].sum() - d["stats"]["sum_e"]) < 1e-9 * abs(d["stats"]["sum_e"])


@pytest.mark.parametrize("omega", [1.0, 2.0])
def test_oscillator_example_zero_variance(omega):
    # examples/custom_operator.rs: Runner::run(1000, 1) of the Gaussian trial
    # under HarmonicHamiltonian, box side 1; every sample is 3 omega / 2
    sj = S.harmonic_oscillator(omega)
    h = M.harmonic_hamiltonian(omega)
    d = device_run(sj, 256, 1000, 1, h=h, side=1.0)
    assert np.all(np.abs(d["el"] - 1.5 * omega) < 1e-12)
    st = d["stats"]
    mean = st["sum_e"] / st["n_samples"]
    assert abs(mean - 1.5 * omega) < 1e-12
    o = oracle_run(sj, 256, 1000, 1, h=h, side=1.0)
    assert np.array_equal(d["masks"], o["accept_mask"])


def test_oscillator_dmc_zero_variance():
    sj = S.harmonic_oscillator(1.0)
    with M.Context([M.slater_jastrow(sj, 128)], seed=1, step=0.02, block_size=10,
                   proposal=M.PROPOSAL_DIFFUSE) as ctx:
        ctx.set_hamiltonian(M.harmonic_hamiltonian(1.0))
        ctx.dmc_init(1.5)
        ctx.dmc_diffuse(0.02, 40, 10, 1)
        steps = ctx.dmc_read_steps()
        w, el = ctx.dmc_read_walkers()
    assert np.all(np.abs(steps - 1.5) < 1e-12) and np.all(np.abs(el - 1.5) < 1e-12)
    assert np.all(np.abs(w - 1.0) < 1e-12)


def test_operator_validation():
    with M.Context([M.h2(16, 1.0, 0.5, 0.3)], seed=0, step=0.5, block_size=5) as ctx:
        with pytest.raises(M.MoleError) as ei:
            ctx.set_hamiltonian(M.electronic_hamiltonian())
        assert ei.value.status == M.ERR_UNSUPPORTED
    with M.Context([M.slater_jastrow(S.atom(3), 16)], seed=0, step=0.5, block_size=5) as ctx:
        with pytest.raises(M.MoleError):
            ctx.set_hamiltonian([(17, 1.0)])
        with pytest.raises(M.MoleError):
            ctx.set_observables([[(M.TERM_RADIAL_POWER, 1.0, 0.5, 0, 0, 0)]])
        with pytest.raises(M.MoleError):
            ctx.set_observables([[(M.TERM_KINETIC, 1.0)]] * (M.MAX_OBSERVABLES + 1))
    with pytest.raises(M.MoleError) as ei:
        M.Context([M.slater_jastrow(S.harmonic_oscillator(1.0), 16)], seed=0, step=0.5, block_size=5,
                  proposal=M.PROPOSAL_LEARNED)
    assert ei.value.status == M.ERR_UNSUPPORTED
