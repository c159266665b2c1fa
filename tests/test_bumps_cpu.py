"""The host restatement of the seeded bumps initial condition, on the CPU:
SplitMix64 against its published outputs (state 0 and state 1234567, the
vectors of Vigna's splitmix64.c and of the SplittableRandom test suites), and
the draws' ranges.  The device initial state is compared with this
restatement by the GPU tests (test_gpu_precision.py)."""
import numpy as np

from bumps_ref import bump_draws, bumps_state, splitmix64


def _outputs(state, n):
    out = []
    for _ in range(n):
        state, z = splitmix64(state)
        out.append(z)
    return out


def test_splitmix64_published_outputs():
    assert _outputs(0, 4) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    assert _outputs(1234567, 5) == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                    4593380528125082431, 16408922859458223821]


def test_bump_draw_ranges():
    d = bump_draws(20251202)
    assert d.shape == (16, 5)
    assert np.all((d[:, :2] >= 0) & (d[:, :2] < 1))
    assert np.all((d[:, 2:4] >= 0.1) & (d[:, 2:4] < 2.0))
    assert np.all((d[:, 4] >= 0.02) & (d[:, 4] < 0.1))
    assert not np.array_equal(d, bump_draws(20251203))


def test_bumps_state_closed_form():
    """Far from every bump the state is the diagonal problem; at a bump centre
    the added density is at least that bump's amplitude."""
    nx = 128
    d = bump_draws(7)
    u = bumps_state(nx, nx, 1.4, 7)
    base = bumps_state(nx, nx, 1.4, 7) * 0
    i = np.arange(nx + 4)[:, None]
    j = np.arange(nx + 4)[None, :]
    x, y = (i - 2) / nx, (j - 2) / nx
    base[..., 0] = np.where(x + y < 1, 1.0, 0.125)
    base[..., 1] = np.where(x + y < 1, 1.0, 0.1) / 0.4
    assert np.all(u[..., 0] >= base[..., 0]) and np.all(u[..., 1] >= base[..., 1] * (1 - 1e-15))
    for cx, cy, ar, ap, sg in d:
        ci, cj = int(cx * nx) + 2, int(cy * nx) + 2
        assert u[ci, cj, 0] - base[ci, cj, 0] >= ar * np.exp(-(1.0 / nx) ** 2 / sg ** 2) * (1 - 1e-12)
