"""rp_fit_phi_noalloc (the basis at arbitrary points without device allocation) against rp_fit_phi.

The no-alloc entry point stages the k-set, shifts, atoms, points and basis rows through a static
per-GPU array and launches the same kernel, so its rows must equal rp_fit_phi's bitwise, on two calls
in a row (the second smaller, so stale rows of the first would show), and the context's device
memory must not grow.
"""
import numpy as np
import pytest

from pyrepeat_amd import _abi, synth

pytestmark = pytest.mark.gpu


def _points(s, m, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, (m, 3)) @ s.box


@pytest.mark.parametrize("name,fit_type", [("crystal200", 1), ("tiny", 0)])
def test_phi_noalloc_equals_phi(name, fit_type):
    s = synth.crystal200() if name == "crystal200" else synth.tiny(seed=4, n_atoms=5, n=12, L=9.0, triclinic=True)
    print(f"no-alloc phi: {name}, fit-type {fit_type}", flush=True)
    ctx = _abi.Context(0, 0)
    f = _abi.Fit(ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, fit_type,
                 _abi.RP_PREC_FP64)
    # two calls in a row, the second smaller; 30 000 points of 200 atoms span three scratch blocks
    for m, seed in ((30000 if name == "crystal200" else 2000, 1), (777, 2)):
        g = _points(s, m, seed)
        before = ctx.hbm_in_use()
        got = f.phi_noalloc(g)
        assert ctx.hbm_in_use() == before
        ref = f.phi(g)
        print(f"  m={m}: max|diff| {np.max(np.abs(got - ref)):.1e}", flush=True)
        np.testing.assert_array_equal(got, ref)
        assert np.all(np.isfinite(got)) and np.max(np.abs(got)) > 0.0
    assert f.phi_noalloc(np.empty((0, 3))).shape == (0, s.n_atoms)
    f.close()
    ctx.close()


def test_phi_noalloc_first_call_on_a_fresh_fit_and_limit():
    """A fit that has not been accumulated: the no-alloc call needs only the fit's host-side Ewald setup and
    adds no device memory; more atoms than the scratch holds is a clear RP_EINVAL."""
    s = synth.tiny(seed=7, n_atoms=4, n=12, L=9.0)
    ctx = _abi.Context(0, 0)
    f = _abi.Fit(ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1,
                 _abi.RP_PREC_OZAKI_REFINED)
    g = _points(s, 300, 3)
    before = ctx.hbm_in_use()
    got = f.phi_noalloc(g)
    assert ctx.hbm_in_use() == before
    np.testing.assert_array_equal(got, f.phi(g))
    f.close()
    # 4097 atoms on a lattice, bare 1/r basis and a small vdW factor so the filter keeps points
    n = 4097
    L = 64.0
    idx = np.arange(n)
    pos = np.stack([idx % 17, (idx // 17) % 17, idx // 289], axis=1) * (L / 17.0) + 0.5
    esp = np.zeros(8 * 8 * 8)
    big = _abi.Fit(ctx, [8, 8, 8], np.eye(3) * (L / 8), np.ones(n, dtype=np.int64), pos, esp, _abi.RP_ORDER_FLAT, 0,
                   _abi.RP_PREC_FP64, vdw_factor=0.01)
    with pytest.raises(_abi.RepeatHipError) as e:
        big.phi_noalloc(np.zeros((4, 3)))
    assert e.value.code == _abi.RP_EINVAL and "4096" in str(e.value)
    big.close()
    ctx.close()
