"""CPU checks of the differentiable-rollout helpers (the losses; the rollout itself is GPU)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _helpers():
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "peclet-dem_amd"))
    from peclet.dem.adjoint import STATE_COLUMNS, bed_height, packing_fraction_below
    return STATE_COLUMNS, bed_height, packing_fraction_below


def test_state_columns_cover_the_13_state_numbers():
    cols, _, _ = _helpers()
    covered = sorted(i for s in cols.values() for i in range(13)[s])
    assert covered == list(range(13))


def test_bed_height_is_the_mean_centre_height_and_its_gradient_is_uniform():
    _, bed_height, _ = _helpers()
    pos = torch.tensor(np.random.default_rng(0).uniform(0, 1, (50, 3)), requires_grad=True)
    h = bed_height(pos)
    assert torch.isclose(h, pos[:, 2].mean())
    h.backward()
    g = pos.grad.numpy()
    assert np.allclose(g[:, 2], 1.0 / 50) and np.all(g[:, :2] == 0.0)


def test_packing_fraction_of_a_uniform_layer():
    # a cubic lattice of spheres (radius r, spacing 2r) filling the bed: phi = pi / 6 in the layer
    _, _, phi_below = _helpers()
    r, m = 0.5, 10
    g = (np.arange(m) + 0.5) * 2 * r
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    pos = torch.tensor(np.stack([x.ravel(), y.ravel(), z.ravel()], 1))
    rad = np.full(len(pos), r)
    area = (m * 2 * r) ** 2
    # H = 2 * mean(z) = m * 2r; the layer below 0.6 H holds the 6 lowest lattice planes
    phi = phi_below(pos, rad, area).item()
    expected = 6 * m * m * (4.0 / 3.0) * np.pi * r ** 3 / (area * 0.6 * m * 2 * r)
    assert phi == pytest.approx(expected, rel=1e-12)


def test_packing_fraction_gradient_flows_through_the_bed_height():
    _, _, phi_below = _helpers()
    rng = np.random.default_rng(1)
    pos = torch.tensor(rng.uniform(0, 1, (40, 3)), requires_grad=True)
    rad = np.full(40, 0.05)
    phi = phi_below(pos, rad, 1.0)
    phi.backward()
    g = pos.grad.numpy()
    # dphi/dz_i = -phi / top * d(top)/dz_i with top = 0.6 * 2 * mean(z): the same for every body
    top = 0.6 * 2.0 * pos[:, 2].detach().numpy().mean()
    assert np.allclose(g[:, 2], -phi.item() / top * (1.2 / 40), rtol=1e-12)
    assert np.all(g[:, :2] == 0.0)
