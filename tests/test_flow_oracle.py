"""Pin the float64 flow reference (CPU): inverse(forward) = identity, the log-determinant
against a brute-force Jacobian determinant, log rho against the PIMC oracle's O-matrix
product, and the whitening against the harmonic covariance."""
import numpy as np
import pytest
import torch

from flow_helpers import random_params, ref_for
from helpers import make_data, oracle_setup, stage_dict_models
from oracle import pimc_oracle as po
from oracle.flow_oracle import FlowRef
from pibronic_amd import synthetic


@pytest.fixture(scope="module")
def small(tmp_path_factory):
    FS = stage_dict_models(tmp_path_factory.mktemp("f"), synthetic.generate(3, 1, quadratic=False, seed=4))
    data = make_data(FS, 300.0, 8, 10, 10)
    return FS, data


def test_inverse_of_forward(small):
    FS, data = small
    ref = ref_for(data, layers=4, hidden=8, smax=2.0)
    flat = random_params(ref, 1, scale=0.3)
    q, _ = data.draw_host_sample(np.random.default_rng(0), 16)
    R0 = torch.as_tensor(q)
    R, ld = ref.forward(R0, flat)
    lq, R0b = ref.log_q(R, flat)
    assert torch.allclose(R0b, R0, atol=1e-10)
    assert torch.allclose(lq, ref.log_rho(R0) - ld, atol=1e-9)


def test_logdet_is_the_jacobian_determinant(small):
    FS, data = small
    ref = ref_for(data, layers=4, hidden=8, smax=2.0)   # D = N*P = 8
    flat = random_params(ref, 2, scale=0.3)
    q, _ = data.draw_host_sample(np.random.default_rng(1), 3)
    for b in range(3):
        R0 = torch.as_tensor(q[b:b + 1]).clone()
        J = torch.autograd.functional.jacobian(lambda x: ref.forward(x, flat)[0].reshape(-1), R0)
        J = J.reshape(ref.D, ref.D)
        _, ld = ref.forward(R0, flat)
        assert torch.linalg.slogdet(J)[1].item() == pytest.approx(ld.item(), abs=1e-9)


def test_log_rho_is_the_oracle_denominator(tmp_path):
    FS = stage_dict_models(tmp_path, synthetic.generate(3, 2, quadratic=False, seed=2))
    data = make_data(FS, 300.0, 6, 10, 10)
    s = oracle_setup(FS, 300.0, 6)
    q, _ = s.draw_q(np.random.default_rng(3), 20)
    o = po.evaluate(s, q, scale=False)
    ref = ref_for(data, 2, 4, 2.0)
    lr = ref.log_rho(torch.as_tensor(q)).numpy()
    assert np.allclose(lr, np.log(o["s_rho"]), rtol=0, atol=1e-11)


def test_whitening_of_a_single_surface_is_standard_normal(tmp_path):
    m = synthetic.generate(1, 2, quadratic=False, seed=3)
    FS = stage_dict_models(tmp_path, m)
    data = make_data(FS, 300.0, 8, 10, 10)
    q, _ = data.draw_host_sample(np.random.default_rng(5), 40000)
    ref = ref_for(data, 1, 4, 2.0)
    z = ref.whiten(torch.as_tensor(q)).numpy()
    assert np.abs(z.mean(axis=0)).max() < 0.05
    assert np.abs(np.cov(z, rowvar=False) - np.eye(ref.D)).max() < 0.05


def test_identity_flow_is_rho(small):
    FS, data = small
    ref = ref_for(data, 2, 4, 2.0)
    flat = np.zeros(ref.L * (ref.H * ref.h + ref.h + ref.h * ref.h + ref.h + ref.h * ref.D + ref.D))
    q, _ = data.draw_host_sample(np.random.default_rng(2), 8)
    lq, R0 = ref.log_q(torch.as_tensor(q), flat)
    assert torch.allclose(R0, torch.as_tensor(q), atol=1e-12)
    assert torch.allclose(lq, ref.log_rho(torch.as_tensor(q)), atol=1e-12)
