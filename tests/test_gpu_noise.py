"""GPU noise trajectories (aleph_hip_run_noisy) against the reference's Aer
fixtures (crates/aleph-sv/tests/noise_oracle.rs: 100k shots, 5 sigma band),
shot by shot against the CPU trajectory restatement (same random stream),
and the driver contracts of crates/aleph-sv/tests/noise_driver.rs."""

import ctypes as C

import numpy as np
import pytest

from aleph_amd import _lib as L
from aleph_amd import noise as N
from aleph_amd.backend import TooManyQubits, _ptr
from aleph_amd.circuit import Circuit
from oracle import noise_oracle as NO
from tests.helpers import (DISTRIBUTION_SHOTS, NOISE_FIXTURES, assert_distribution_close, load_noise_fixture,
                           mixed_noise_case, noise_fixture_case)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", NOISE_FIXTURES)
def test_aer_noise_fixtures(hip, name):
    c, m = noise_fixture_case(name)
    counts = N.run_noisy(c, m, DISTRIBUTION_SHOTS, 0, backend=hip)
    assert counts.sum() == DISTRIBUTION_SHOTS
    assert_distribution_close(counts, np.asarray(load_noise_fixture(name)["exact_probs"]), DISTRIBUTION_SHOTS, name)


@pytest.mark.parametrize("n,depth,shots", [(5, 4, 400), (9, 3, 200), (14, 2, 24), (15, 1, 12)])
def test_trajectories_equal_cpu_restatement_shot_by_shot(hip, oracle_lib, n, depth, shots):
    # Up to 13 qubits the state lives in LDS, beyond in device slices.
    c, m = mixed_noise_case(n, depth, seed=n)
    prog, att = N.resolve(c, m)
    want = NO.trajectory_outcomes(prog, att, m.readout_map(), n, 99, shots, shot0=1000)
    got = N.outcomes(c, m, shots, 99, backend=hip, shot_offset=1000)
    assert np.array_equal(got, want), f"{int(np.sum(got != want))} of {shots} shots differ"


def test_mixed_model_matches_density_matrix(hip, oracle_lib):
    c, m = mixed_noise_case(5, 3, seed=8)
    prog, att = N.resolve(c, m)
    exact = NO.exact_distribution(prog, att, m.readout_map(), 5)
    counts = N.run_noisy(c, m, DISTRIBUTION_SHOTS, 21, backend=hip)
    assert_distribution_close(counts, exact, DISTRIBUTION_SHOTS, "mixed n=5")


def test_shots_split_over_calls_equal_one_call(hip):
    c, m = mixed_noise_case(6, 3, seed=2)
    one = N.outcomes(c, m, 1000, 7, backend=hip)
    parts = np.concatenate([N.outcomes(c, m, 300, 7, backend=hip, shot_offset=0),
                            N.outcomes(c, m, 700, 7, backend=hip, shot_offset=300)])
    assert np.array_equal(one, parts)
    assert not np.array_equal(one, N.outcomes(c, m, 1000, 8, backend=hip))
    sparse = N.run_noisy_sparse(c, m, 1000, 7, backend=hip)
    dense = N.run_noisy(c, m, 1000, 7, backend=hip)
    assert sparse == [(i, int(v)) for i, v in enumerate(dense) if v]


def test_noise_fires_only_under_the_gate_name(hip):
    # noise_driver.rs:44-66
    c = Circuit(1).x(0)
    m = N.NoiseModel()
    m.add_all_qubit_quantum_error(N.depolarizing_error(0.3, 1), ["X"])
    counts = N.run_noisy(c, m, 50_000, 11, backend=hip)
    assert counts[0] > 1000
    wrong = N.NoiseModel()
    wrong.add_all_qubit_quantum_error(N.depolarizing_error(0.3, 1), ["x"])
    counts = N.run_noisy(c, wrong, 50_000, 11, backend=hip)
    assert counts[0] == 0 and counts[1] == 50_000


def test_empty_model_reproduces_noiseless_distribution(hip):
    # noise_driver.rs:68-75
    counts = N.run_noisy(Circuit(2).h(0).cx(0, 1), N.NoiseModel(), DISTRIBUTION_SHOTS, 1, backend=hip)
    assert_distribution_close(counts, np.array([0.5, 0.0, 0.0, 0.5]), DISTRIBUTION_SHOTS, "noise_empty_bell")


def test_zero_shots_and_limits(hip):
    c, m = mixed_noise_case(4, 1, seed=1)
    assert N.run_noisy(c, m, 0, 1, backend=hip).sum() == 0 and N.run_noisy_sparse(c, m, 0, 1, backend=hip) == []
    with pytest.raises(TooManyQubits):
        N.outcomes(Circuit(25).h(0), N.NoiseModel(), 4, 1, backend=hip)


def test_library_rejects_malformed_channels(hip):
    """The library validates channels itself (a caller that skips the host
    checks still gets the reference's InvalidChannel errors)."""
    c = Circuit(2).h(0).cx(0, 1)
    words, reals = c.encode()
    w = np.ascontiguousarray([0.5, 0.4])
    codes = np.ascontiguousarray([0, 1], dtype=np.uint8)
    ch = (L.Channel * 1)()
    ch[0].after, ch[0].kind, ch[0].n_qubits, ch[0].n_branches = 0, L.CHANNEL_PAULI, 1, 2
    ch[0].weights, ch[0].paulis = _ptr(w, C.c_double), _ptr(codes, C.c_uint8)
    out = np.zeros(4, dtype=np.uint64)
    ro = np.zeros(8)
    lib = L.load()

    def run(channels, count, mask=0):
        return lib.aleph_hip_run_noisy(hip._ctx, 2, _ptr(words, C.c_uint32), words.size, _ptr(reals, C.c_double),
                                       reals.size, channels, count, _ptr(ro, C.c_double), C.c_uint64(mask), 4,
                                       C.c_uint64(1), C.c_uint64(0), 0, _ptr(out, C.c_uint64))

    def kind():
        e = L.Error()
        lib.aleph_hip_last_error(hip._ctx, C.byref(e))
        return e.kind.decode(), e.message.decode()

    assert run(ch, 1) == L.ERR_INVALID_ARGUMENT and kind() == ("invalid-channel", "Pauli probabilities must sum to one")
    w[1] = 0.5
    ch[0].after = 5
    assert run(ch, 1) == L.ERR_INVALID_ARGUMENT and "past the end" in kind()[1]
    ch[0].after = 0
    ro[:4] = [0.9, 0.2, 0.0, 1.0]
    assert run(ch, 1, mask=1) == L.ERR_INVALID_ARGUMENT and "readout" in kind()[1]
    ro[:4] = [0.9, 0.1, 0.0, 1.0]
    assert run(ch, 1, mask=1) == L.OK
