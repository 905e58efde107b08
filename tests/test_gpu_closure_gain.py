"""What the trained closure network gains over the method without it (Smagorinsky, the zero network), a priori.

At a quarter of the ground boundary-layer workload (coarse 128 x 32 x 64; fine 256 x 64 x 128), the network is trained
with the matrix-core kernels on the filtered 2x-resolution field of one building and scored on the filtered field of
another building it has not seen: the weighted least-squares loss of the sub-filter stress (the training objective)
must be at least 10% below the zero network's, on the held-out samples as on the training samples. The gain is a
rescaling of the Smagorinsky viscosity (the stress invariants do not predict the per-sample target, so the trained
output is nearly constant); a posteriori it does not halve the coarse-grid Cd error of the workload's building, which
the closure workflow measures on the full-size protocol.
"""
import numpy as np
import pytest

import aerosim_amd as A
from aerosim_amd import closure as C

pytestmark = pytest.mark.gpu

BASE = C.BuildingCase().scaled(0.25)  # coarse 128 x 32 x 64, building 10 x 5 x 15 cells


def _samples(size, seed, steps=3000, snapshots=12):
    """Filtered samples of the 2x-resolution plain run of a building: snapshots spread over the second half."""
    cl = A.Closure()
    ctx, _, _ = C.sized(BASE, size).scaled(2).context(seed=seed)
    t = 0
    for k in range(snapshots):
        e = steps // 2 + (steps // 2) * (k + 1) // snapshots
        ctx.step(e - t, readout=True)
        t = e
        cl.sample_fine(ctx, 2, 100 * seed + k, BASE.nu, BASE.cs)
    ctx.close()
    X, T = cl.samples()
    cl.close()
    return X, T


def test_trained_network_beats_smagorinsky_a_priori_on_a_held_out_building():
    Xa, Ta = _samples((10, 6, 14), 1)  # training building
    Xb, Tb = _samples((12, 4, 16), 2)  # held out
    assert len(Ta) > 100000 and len(Tb) > 100000
    net = A.Closure()  # starts as the zero network: exactly the Smagorinsky closure
    perm = np.random.default_rng(0).permutation(len(Ta))
    net.add_samples(Xa[perm], Ta[perm])
    smag_train, smag_held = net.grad(Xa, Ta)[0], net.grad(Xb, Tb)[0]
    net.train(1000, batch=4096, lr=2e-3)
    net_train, net_held = net.grad(Xa, Ta)[0], net.grad(Xb, Tb)[0]
    y = net.predict(Xb)
    factor = float(np.mean(np.log1p(np.exp(y + np.log(np.e - 1.0)))))
    print(f"loss, training building: Smagorinsky {smag_train:.5g}, trained {net_train:.5g} "
          f"({(net_train / smag_train - 1) * 100:+.1f}%); held-out building: Smagorinsky {smag_held:.5g}, trained "
          f"{net_held:.5g} ({(net_held / smag_held - 1) * 100:+.1f}%); mean factor on the held-out samples {factor:.3f}")
    net.close()
    assert net_train <= 0.9 * smag_train
    assert net_held <= 0.9 * smag_held
