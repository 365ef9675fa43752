"""Host re-gridding (cdc_resample_state, no GPU): the checkpoint up-sampler's
piecewise-constant injection (upsampler/src/main.rs:27-72) and its
conservative inverse.  Pinned by: injection equals numpy's repeat of each
block's cells (the up-sampler's conserved[i][j] = old[i/2][j/2]); down(up(u))
is u exactly; the area average conserves the total of every field; an
area average checked against numpy's reshape-mean on random data; the ist /
oec rules (copied with num_blocks kept, summed over merged blocks)."""
import numpy as np
import pytest

from cdc_amd import Mesh, resample_state


def rand_state(nb, bs, nv, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.5, 2.0, (nb * nb, bs, bs, nv))
    ist = rng.normal(size=(nb * nb, 7, nv))
    oec = rng.normal(size=(nb * nb, 7, 4))
    return u, ist, oec


def to_global(u, nb, bs):
    nv = u.shape[-1]
    return u.reshape(nb, nb, bs, bs, nv).transpose(0, 2, 1, 3, 4).reshape(nb * bs, nb * bs, nv)


@pytest.mark.parametrize("nv", [3, 4])
def test_upsampler_is_blockwise_injection(nv):
    nb, bs = 3, 8
    u, ist, oec = rand_state(nb, bs, nv, 1)
    up, ui, ue = resample_state(Mesh(nb, bs, 4.0), Mesh(nb, 2 * bs, 4.0), u, ist, oec)
    want = np.repeat(np.repeat(u, 2, axis=1), 2, axis=2)
    np.testing.assert_array_equal(up, want)
    np.testing.assert_array_equal(ui, ist)
    np.testing.assert_array_equal(ue, oec)


@pytest.mark.parametrize("ratio", [2, 4, 8])
def test_down_of_up_is_identity(ratio):
    nb, bs = 2, 8
    u, ist, oec = rand_state(nb, bs, 4, ratio)
    fine = Mesh(nb, bs * ratio, 4.0)
    up, _, _ = resample_state(Mesh(nb, bs, 4.0), fine, u)
    back, bi, be = resample_state(fine, Mesh(nb, bs, 4.0), up, ist, oec)
    np.testing.assert_array_equal(back, u)
    np.testing.assert_array_equal(bi, ist)


def test_reblocked_average_matches_numpy_and_conserves():
    # 2048^2-style re-blocking at small size: 8x8 blocks of 16^2 -> 2x2 blocks of 16^2 (ratio 4)
    src, dst = Mesh(8, 16, 8.0), Mesh(2, 16, 8.0)
    u, ist, oec = rand_state(8, 16, 4, 5)
    d, di, de = resample_state(src, dst, u, ist, oec)
    g = to_global(u, 8, 16)
    want = g.reshape(32, 4, 32, 4, 4).mean(axis=(1, 3))
    got = to_global(d, 2, 16)
    np.testing.assert_allclose(got, want, rtol=2e-16 * 16, atol=0)
    # totals of every field (area-weighted) agree to rounding
    np.testing.assert_allclose(got.sum(axis=(0, 1)) * 16, g.sum(axis=(0, 1)), rtol=1e-14)
    # each coarse block is a 4x4 group of fine blocks: ist / oec summed
    want_i = ist.reshape(2, 4, 2, 4, 7, 4).sum(axis=(1, 3)).reshape(4, 7, 4)
    np.testing.assert_allclose(di, want_i, rtol=1e-14, atol=1e-14)
    want_e = oec.reshape(2, 4, 2, 4, 7, 4).sum(axis=(1, 3)).reshape(4, 7, 4)
    np.testing.assert_allclose(de, want_e, rtol=1e-14, atol=1e-14)


def test_rejects_non_power_of_two_and_other_domains():
    u, _, _ = rand_state(2, 8, 3, 0)
    with pytest.raises(Exception):
        resample_state(Mesh(2, 8, 4.0), Mesh(2, 24, 4.0), u)
    with pytest.raises(Exception):
        resample_state(Mesh(2, 8, 4.0), Mesh(2, 16, 5.0), u)
